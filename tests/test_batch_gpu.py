"""GPU parity of the batched uniform solve (sg_batch_*): many independent periodic problems of
uneven sizes, n = 2^a or 3 * 2^a, solved together.

* FP64 batch: every problem bit for bit the CPU oracle's mg_cycle (which extends the reference's
  level sequence to n = 3 * 2^a with an exact 3^3 coarsest solve, pinned against the FFT).
* Mixed batch: a power-of-two problem bit for bit the single-grid mixed solver (same fused
  kernels, same FP64 outer step, either coarse-grid solve); every problem reaches 1e-10 and stays within 1e-9 (rel. L2) of
  the oracle's converged FP64 solution; a problem's result does not depend on the other problems
  of the batch (it equals the batch-of-one solve), nor on graph replay.
* The 48-problem uneven batch (sizes 32..192): all problems converge to 1e-10 in mixed precision;
  mixed vs FP64 batch and a 192^3 problem against the FFT solution."""
import numpy as np
import pytest
import torch

import oracle
from oracle import fft_poisson as fp
from code_only_amd import problems
from code_only_amd.batch import SG_COARSE_EXACT, SG_COARSE_VCYCLE, SG_PREC_FP64, SG_PREC_MIXED, BatchSolver
from code_only_amd.uniform import UniformSolver

pytestmark = pytest.mark.gpu


def rhs_list(sizes, seed=0):
    out = []
    for q, n in enumerate(sizes):
        sig = 4.0 if n >= 32 else 1.0
        out.append(problems.clumpy_rhs(n, seed=seed + q, nclump=1 + (q * 5) % 16, sigma_cells=sig))
    return out


def demean(a):
    return a - a.mean()


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


SMALL = [2, 3, 4, 6, 12, 16, 24, 32, 48, 96, 64]


def test_fp64_batch_bitwise_vs_oracle(gpu):
    srcs = rhs_list(SMALL, seed=1)
    b = BatchSolver(SMALL, precisions=SG_PREC_FP64)
    psis, st = b.solve_host(srcs, tol=1e-10, maxit=30, precision=SG_PREC_FP64)
    for q, n in enumerate(SMALL):
        po, so = oracle.mg_cycle(srcs[q], 1.0 / n, tol=1e-10, maxit=30)
        print(n, so["iterations"], st[q]["iterations"], so["rel_linf"], st[q]["rel_linf"], flush=True)
        assert st[q]["status"] == 0 and so["status"] == 0
        assert st[q]["iterations"] == so["iterations"]
        assert st[q]["rel_linf"] == so["rel_linf"]
        assert np.array_equal(psis[q], po), f"n = {n}: max |diff| {np.abs(psis[q] - po).max():.3e}"
    b.close()


def test_fp64_batch_z_marching_outer_step_bitwise_vs_oracle(gpu):
    """When every problem's n is a multiple of 32 the outer update + residual runs as z-marching
    patches; the FP64 batch still equals the oracle bit for bit."""
    sizes = [32, 96, 64]
    srcs = rhs_list(sizes, seed=4)
    b = BatchSolver(sizes, precisions=SG_PREC_FP64)
    psis, st = b.solve_host(srcs, tol=1e-10, maxit=30, precision=SG_PREC_FP64)
    for q, n in enumerate(sizes):
        po, so = oracle.mg_cycle(srcs[q], 1.0 / n, tol=1e-10, maxit=30)
        assert st[q]["status"] == 0 and st[q]["iterations"] == so["iterations"]
        assert st[q]["rel_linf"] == so["rel_linf"]
        assert np.array_equal(psis[q], po), f"n = {n}"
    b.close()


@pytest.mark.parametrize("coarse", [SG_COARSE_EXACT, SG_COARSE_VCYCLE])
def test_mixed_batch_power_of_two_equals_single_grid_solver(gpu, coarse):
    sizes = [64, 16, 32, 128, 8]
    srcs = rhs_list(sizes, seed=5)
    b = BatchSolver(sizes, precisions=SG_PREC_MIXED)
    psis, st = b.solve_host(srcs, tol=1e-10, maxit=30, precision=SG_PREC_MIXED, coarse=coarse)
    for q, n in enumerate(sizes):
        s = UniformSolver(n, precisions=SG_PREC_MIXED)
        p1 = torch.zeros((n,) * 3, dtype=torch.float64, device=gpu)
        s1 = s.cycle(p1, torch.tensor(srcs[q], device=gpu), tol=1e-10, maxit=30, precision=SG_PREC_MIXED,
                     coarse=coarse)
        s.close()
        print(n, s1["iterations"], st[q]["iterations"], s1["rel_linf"], st[q]["rel_linf"], flush=True)
        assert st[q]["iterations"] == s1["iterations"] and st[q]["rel_linf"] == s1["rel_linf"]
        assert np.array_equal(psis[q], p1.cpu().numpy()), f"n = {n}"
    b.close()


@pytest.mark.parametrize("graph", [False, True])
def test_mixed_batch_against_oracle(gpu, graph):
    sizes = [24, 48, 96, 12, 6, 3, 32, 192]
    srcs = rhs_list(sizes, seed=9)
    b = BatchSolver(sizes, precisions=SG_PREC_MIXED)
    psis, st = b.solve_host(srcs, tol=1e-10, maxit=30, precision=SG_PREC_MIXED, use_graph=graph)
    for q, n in enumerate(sizes):
        po, so = oracle.mg_cycle(srcs[q], 1.0 / n, tol=1e-12, maxit=40)
        e = rel(demean(psis[q]), demean(po))
        print(n, st[q]["iterations"], st[q]["rel_linf"], "err vs oracle", e, flush=True)
        assert st[q]["status"] == 0 and st[q]["rel_linf"] < 1e-10
        assert e < 1e-9
    b.close()


def test_batch_results_independent_of_batch_and_graph(gpu):
    """Each problem of a mixed batch equals its batch-of-one solve bit for bit (problems that
    converge early drop out without disturbing the others), graph or eager, on repeated calls."""
    sizes = [48, 3, 96, 12, 24, 32]
    srcs = rhs_list(sizes, seed=9)
    b = BatchSolver(sizes, precisions=SG_PREC_MIXED)
    pg, sg_ = b.solve_host(srcs, tol=1e-10, maxit=30, use_graph=True)
    pg2, _ = b.solve_host(srcs, tol=1e-10, maxit=30, use_graph=True)
    pe, se = b.solve_host(srcs, tol=1e-10, maxit=30, use_graph=False)
    its = [s["iterations"] for s in sg_]
    print("iterations", its, flush=True)
    assert len(set(its)) >= 2, "the problems should converge after different numbers of iterations"
    for q, n in enumerate(sizes):
        assert np.array_equal(pg[q], pe[q]) and np.array_equal(pg[q], pg2[q])
        one = BatchSolver([n], precisions=SG_PREC_MIXED)
        p1, s1 = one.solve_host([srcs[q]], tol=1e-10, maxit=30)
        one.close()
        assert s1[0]["iterations"] == its[q]
        assert np.array_equal(p1[0], pg[q]), f"n = {n}"
    b.close()


def test_batch_rejects_bad_sizes(gpu):
    from code_only_amd._lib import SgError
    for sizes in ([80], [1], [32, 40]):
        with pytest.raises(SgError):
            BatchSolver(sizes)


def test_uneven_48_problem_batch(gpu):
    """The full uneven batch (sizes among 32, 64, 96, 128, 192; 1..16 clumps each)."""
    ns, rhs = problems.batch_problems(48, seed=0, device=gpu)
    b = BatchSolver(ns)
    pm = [torch.zeros_like(r) for r in rhs]
    p64 = [torch.zeros_like(r) for r in rhs]
    sm = b.solve(pm, rhs, tol=1e-10, maxit=30, precision=SG_PREC_MIXED)
    s64 = b.solve(p64, rhs, tol=1e-10, maxit=30, precision=SG_PREC_FP64)
    print("cells", sum(n ** 3 for n in ns), "mixed ms", sm[0]["ms"], "fp64 ms", s64[0]["ms"],
          "iterations mixed", sorted(set(s["iterations"] for s in sm)),
          "fp64", sorted(set(s["iterations"] for s in s64)), flush=True)
    worst = 0.0
    for q in range(len(ns)):
        assert sm[q]["status"] == 0 and sm[q]["rel_linf"] < 1e-10
        assert s64[q]["status"] == 0 and s64[q]["rel_linf"] < 1e-10
        a, c = pm[q] - pm[q].mean(), p64[q] - p64[q].mean()
        worst = max(worst, (torch.linalg.norm(a - c) / torch.linalg.norm(c)).item())
    print("worst mixed vs fp64 rel L2", worst, flush=True)
    assert worst < 1e-9
    q = ns.index(192)
    ref = fp.solve(rhs[q].cpu().numpy(), 1.0 / 192)
    e = rel(demean(pm[q].cpu().numpy()), ref)
    print("192^3 mixed vs FFT", e, flush=True)
    assert e < 1e-9
    b.close()
