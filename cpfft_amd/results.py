"""Result files of the load steps a deck lists under `output results steps`
(ouresult.f:12-43, called after the roll-over of each such step,
FFT_nr3.f:183), in the reference's flat text form (its output options are
fixed there: flat_file, text_file, ouresult.f:80-91 and 166-173):

  wnd<step>_text   nodal displacements u of the (N+1)^3 brick-mesh nodes,
                   three values per line (format 3e15.6, ouresult.f:107-120)
                   -- cpfft_f2disp's u;
  wee<step>_text   element strains: per voxel the six components of the
                   accumulated unrotated strain eps (CPFFT_EPS), padded with
                   zeros to 7 + 15 = 22 values (ouresult.f:163-166, 232-243)
  wes<step>_text   element stresses: per voxel the six components of the
                   unrotated Cauchy stress (state words CPFFT_W_SIG), padded
                   to 11 + 15 = 26 values (ouresult.f:214-227)

<step> is the step number as i5.5 (less 99999 above 99999,
ouresult.f:531, 622).  Each file starts with the header of
ouddpa_flat_header (ouresult.f:336-397).  Values below 1e-30 in magnitude
are written as zero (ouresult.f:104, 271).  The header's node count is
(N+1)^3 (the reference takes it from `output model`, oumodel.f:78-81, 146).
Element order is the voxel order, x slowest, as the reference's element
numbering.
"""
import os
import time

import numpy as np

from . import EPS, STATE_N

SMALL = 1.0e-30
N_STRAIN_VALUES = 7 + 15   # num_short_strain + 15 (ouresult.f:163-166)
N_STRESS_VALUES = 11 + 15  # num_short_stress + 15


def fortran_e(x, w=15, d=6):
    """One value in the Fortran edit descriptor Ew.d: 0.<d digits>E+ee,
    right-justified in w columns (a three-digit exponent drops the E)."""
    x = float(x)
    if x == 0.0:
        body = "0." + "0" * d + "E+00"
    else:
        m, e = f"{abs(x):.{d - 1}e}".split("e")  # d significant digits, correctly rounded
        digits = m.replace(".", "")
        ex = int(e) + 1
        sign = "-" if ex < 0 else "+"
        exp = f"E{sign}{abs(ex):02d}" if abs(ex) < 100 else f"{sign}{abs(ex):03d}"
        body = ("-" if x < 0 else "") + "0." + digits + exp
    return body.rjust(w)


def _zero_small(a):
    a = np.array(a, dtype=np.float64, copy=True)
    a[np.abs(a) < SMALL] = 0.0
    return a


def _header(kind, quantity, stname, nonode, noelem, step):
    """ouddpa_flat_header: kind 1 nodal vector, 3 element tensor"""
    what = {1: "#  WARP3D nodal results: ", 3: "#  WARP3D element results: "}[kind]
    return "\n".join([
        "#",
        what + quantity.ljust(15)[:15],
        "#  Structure name: " + stname.ljust(8)[:8],
        "#  Model nodes, elements: " + f"{nonode:8d}{noelem:8d}",
        "#  " + time.ctime().ljust(24)[:24],
        "#  Load(time) step: " + f"{step:8d}",
        "#",
    ]) + "\n"


def file_name(prefix, step):
    """wnd / wee / wes + i5.5 step + _text"""
    if step > 99999:
        step -= 99999
    return f"{prefix}{step:05d}_text"


def _rows(values):
    return "".join("".join(fortran_e(v) for v in row) + "\n" for row in values)


def write_results(solver, step, stname, lengths, directory="."):
    """The three result files of load step `step` from the solver's state
    after its roll-over; lengths = the deck's box sizes (l_x, l_y, l_z) for
    the displacements.  Returns the paths written."""
    N = solver.N
    V = N ** 3
    nonode = (N + 1) ** 3
    paths = []
    u = _zero_small(solver.f2disp(lengths=lengths)).reshape(nonode, 3)
    p = os.path.join(directory, file_name("wnd", step))
    with open(p, "w") as f:
        f.write(_header(1, "displacements", stname, nonode, V, step))
        f.write(_rows(u))
    paths.append(p)
    for prefix, quantity, n, comps in (
            ("wee", "strains", N_STRAIN_VALUES, solver.download(EPS)),
            ("wes", "stresses", N_STRESS_VALUES, solver.download(STATE_N)[0:6])):
        vals = np.zeros((V, n))
        vals[:, 0:6] = _zero_small(comps).T
        p = os.path.join(directory, file_name(prefix, step))
        with open(p, "w") as f:
            f.write(_header(3, quantity, stname, nonode, V, step))
            f.write(_rows(vals))
        paths.append(p)
    return paths


def run_deck(path, directory=".", gamma_mode=0, solver_out=None):
    """The deck's `compute`: the load steps through the Fortran driver, one
    step per call, with the result files of the steps `output results`
    lists written after their roll-over (FFT_nr3.f:183).  Returns (rc,
    Pbar [nstep][9], counters [nstep][4], paths written).  solver_out, a
    list, receives the Solver (left open) instead of its being closed."""
    from . import Solver
    from .deck import read_deck
    d = read_deck(path)
    BC, isNBC = d.bc()
    s = Solver(d.N, d.angles(), d.material(), dt=d.dt, gamma_mode=gamma_mode)
    Ps, cnts, paths = [], [], []
    rc = 0
    for k in range(BC.shape[0]):
        rc, P, cnt, _ = s.compute(BC[k:k + 1], isNBC, d.tolPCG, d.tolNR, d.maxIter,
                                  first_step=k + 1)
        if rc != 0:
            break
        Ps.append(P)
        cnts.append(cnt)
        if k + 1 in d.out_steps:
            paths += write_results(s, k + 1, d.stname, d.sizes, directory)
    if solver_out is not None:
        solver_out.append(s)
    else:
        s.close()
    return (rc, np.vstack(Ps) if Ps else np.zeros((0, 9)),
            np.vstack(cnts) if cnts else np.zeros((0, 4), np.int32), paths)
