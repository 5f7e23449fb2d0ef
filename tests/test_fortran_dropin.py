"""The Fortran drop-in modules, driven by Fortran programs through the ISO_C_BINDING shim:
* `mg` (code_only_amd/fortran/mg_gpu.F90) in the style of the reference's mg_test
  (fmg_test.F90:8-77);
* `fmg` (code_only_amd/fortran/fmg_gpu.F90): argument-less fmg_poisson over a hierarchy held in
  a host module `grid` with the reference's names and layout (tests/fortran/grid_host.F90), rho in
  the U1 slot, psi and g written back into the block components.  FP64 results are bitwise the
  oracle's fmg_cycle in both stopping modes."""
import os
import subprocess

import numpy as np
import pytest

import oracle
from oracle import fft_poisson as fp
from code_only_amd import problems

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "code_only_amd", "fortran", "_build", "mg_drop_in_test")
MGDATA_EXE = os.path.join(ROOT, "code_only_amd", "fortran", "_build", "mg_data_drop_in_test")
FMG_EXE = os.path.join(ROOT, "code_only_amd", "fortran", "_build", "fmg_drop_in_test")


def test_driver_is_built_and_linked():
    if not (os.path.exists(EXE) and os.path.exists(FMG_EXE)):
        from code_only_amd import build as b
        b.build_fortran()
    for exe in (EXE, FMG_EXE, MGDATA_EXE):
        out = subprocess.run(["ldd", exe], capture_output=True, text=True, check=True).stdout
        assert "libsg.so" in out


def run_driver(tmp_path, n, prec, f):
    fin, fout = tmp_path / "in.bin", tmp_path / "out.bin"
    f.astype(np.float64).tofile(fin)
    r = subprocess.run([EXE, str(n), str(prec), str(fin), str(fout)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    return np.fromfile(fout, dtype=np.float64).reshape(n, n, n), r.stdout


@pytest.mark.gpu
def test_fortran_mg_cycle_fp64_bitwise(tmp_path):
    n = 32
    f = problems.clumpy_rhs(n, seed=21)
    psi, log = run_driver(tmp_path, n, 1, f)
    ref, st = oracle.mg_cycle(f, 1.0 / n, tol=1e-10, maxit=30)
    assert f"iterations {st['iterations']}" in log
    assert np.array_equal(psi, ref)


@pytest.mark.gpu
def test_fortran_mg_cycle_mixed(tmp_path):
    n = 32
    h = 1.0 / n
    f = problems.clumpy_rhs(n, seed=22)
    psi, _ = run_driver(tmp_path, n, 2, f)
    r = f - fp.apply_laplacian(psi, h)
    assert np.max(np.abs(r)) / np.max(np.abs(f)) < 1e-10


# ---- module mg_data: the reference callers' exchange sequences -----------------------------

@pytest.mark.gpu
@pytest.mark.parametrize("fmglev", [0, 1])
def test_fortran_mg_data_exchange_fp64_bitwise(tmp_path, fmglev):
    """mg_test's sequence (mg_init, mg_dataPrepare((/IRHO, IU/), (/IRHO, IU/)), mg_fill0, mg_cycle,
    mg_dataRestore, fmg_test.F90:38-55) and vmg_slvsml's (mg_dataPrepare((/ISRC, IPSI/), (/jrhs, ju/)),
    mg_cycle(IPSI, ISRC), mg_dataRestore((/ju/), (/IPSI/)), vmg.F90:407-428, here from a non-zero
    guess), over base blocks held in the host's grid / fmg_data modules: both solutions, read back
    from the FMG blocks, are bit for bit the oracle's mg_cycle on the gathered grid."""
    nb = 4
    n = nb * 8 >> fmglev
    h = 2.0 ** fmglev / (nb * 8)
    f = problems.clumpy_rhs(n, seed=23)
    guess = np.random.default_rng(5).standard_normal((n, n, n)) * 1e-3
    fin, fg, o1, o2 = (tmp_path / s for s in ("in.bin", "guess.bin", "o1.bin", "o2.bin"))
    f.astype(np.float64).tofile(fin)
    guess.astype(np.float64).tofile(fg)
    r = subprocess.run([MGDATA_EXE, str(nb), str(fmglev), "1", str(fin), str(fg), str(o1), str(o2)],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    print(r.stdout, flush=True)
    u1 = np.fromfile(o1, dtype=np.float64).reshape(n, n, n)
    u2 = np.fromfile(o2, dtype=np.float64).reshape(n, n, n)
    ref1, s1 = oracle.mg_cycle(f, h, tol=1e-10, maxit=30)
    ref2, s2 = oracle.mg_cycle(f, h, psi=guess, tol=1e-10, maxit=30)
    assert f"mg_test: iterations {s1['iterations']}" in r.stdout
    assert f"vmg_slvsml: iterations {s2['iterations']}" in r.stdout
    assert np.array_equal(u1, ref1) and np.array_equal(u2, ref2)


# ---- module fmg ----------------------------------------------------------------------------

from code_only_amd import amr_hierarchy as ah  # noqa: E402

PI4G = 2.5


def write_fmg_input(path, h, rho, psi0):
    nb = h.nblk
    same = np.where(h.level[h.nbr.clip(0)] == h.level[:, None], h.nbr, -1).astype(np.int32)
    same[h.nbr < 0] = -1
    with open(path, "wb") as fh:
        np.array([nb, h.nlev, h.bc], np.int32).tofile(fh)
        np.array([h.h0, PI4G, *h.x0], np.float64).tofile(fh)
        for a in (h.level, h.ijk, h.parent, h.child, same):
            np.ascontiguousarray(a, np.int32).tofile(fh)
        np.ascontiguousarray(rho, np.float64).tofile(fh)
        np.ascontiguousarray(psi0, np.float64).tofile(fh)


def run_fmg(tmp_path, h, rho, psi0, tol, maxit, stop_mode, prec):
    fin, fout = tmp_path / "fmg_in.bin", tmp_path / "fmg_out.bin"
    write_fmg_input(fin, h, rho, psi0)
    r = subprocess.run([FMG_EXE, str(fin), str(fout), repr(tol), str(maxit), str(stop_mode), str(prec)],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    out = np.fromfile(fout, dtype=np.float64)
    nb = h.nblk
    return out[:nb * 512].reshape(nb, 8, 8, 8), out[nb * 512:].reshape(nb, 3, 8, 8, 8), r.stdout


def fmg_case(name):
    if name == "periodic":
        return ah.random_hierarchy(4, 3, 0.35, seed=7, bc=ah.PERIODIC)
    return ah.nested(base=4, levels=3, bc=ah.SPHERE)


def test_fmg_input_layout_is_readable(tmp_path):
    h = fmg_case("sphere")
    write_fmg_input(tmp_path / "x.bin", h, np.zeros((h.nblk, 8, 8, 8)), np.zeros((h.nblk, 8, 8, 8)))
    hdr = np.fromfile(tmp_path / "x.bin", dtype=np.int32, count=3)
    assert hdr.tolist() == [h.nblk, h.nlev, h.bc]


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["periodic", "sphere"])
@pytest.mark.parametrize("mode", [oracle.STOP_TRUE_RESIDUAL, oracle.STOP_REFERENCE])
def test_fortran_fmg_poisson_fp64_bitwise(tmp_path, name, mode):
    h = fmg_case(name)
    src0 = problems.amr_source(h, seed=31)
    rho = src0 / PI4G
    src = rho * PI4G
    psi0 = np.random.default_rng(3).standard_normal(src.shape) * 1e-3
    tol, maxit = (1e-10, 30) if mode == oracle.STOP_TRUE_RESIDUAL else (1e-3, 10)
    psi, g, log = run_fmg(tmp_path, h, rho, psi0, tol, maxit, mode, 1)
    o = oracle.Amr(h)
    po, go_, so = o.solve(src, psi=psi0, tol=tol, maxit=maxit, mode=mode)
    o.close()
    assert f"iterations {so['iterations']} " in log
    assert np.array_equal(psi, po)
    assert np.array_equal(g, go_)


@pytest.mark.gpu
def test_fortran_fmg_poisson_mixed(tmp_path):
    h = fmg_case("sphere")
    rho = problems.amr_source(h, seed=32) / PI4G
    psi, g, log = run_fmg(tmp_path, h, rho, np.zeros_like(rho), 1e-10, 30, oracle.STOP_TRUE_RESIDUAL, 2)
    o = oracle.Amr(h)
    po, go_, so = o.solve(rho * PI4G, tol=1e-12, maxit=40)
    o.close()
    assert np.linalg.norm(psi - po) / np.linalg.norm(po) < 1e-9
    assert np.linalg.norm(g - go_) / np.linalg.norm(go_) < 1e-8
