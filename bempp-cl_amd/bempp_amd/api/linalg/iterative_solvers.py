"""Iterative solvers on the device (bempp_cl/api/linalg/iterative_solvers.py:10-266).

The reference hands the weak form (a host LinearOperator) to scipy.sparse.linalg.gmres / cg.
Here the weak form is a dense device matrix, and the Krylov iteration runs on the device with
the library's HIP product (FP64 accumulation) as the operator; only the small Hessenberg least
squares problem (restart x restart) is solved on the host. The iteration restates scipy's
algorithms (scipy/sparse/linalg/_isolve/iterative.py, gmres and cg, scipy 1.12+): restarted
GMRES with Givens rotations, the same inner/outer tolerance logic (rtol relative to ||b||,
atol = 0) and the same callback values, so that the iteration counts, residual histories and
iterates agree with the reference to rounding. The Arnoldi step orthogonalises with two
classical Gram-Schmidt passes where scipy uses modified Gram-Schmidt: the Krylov basis is the
same to rounding. On the GPU the step is one library call (bem_gmres_cgs2: three passes over
the basis and the next basis vector, one host copy of the coefficients) instead of 2 (restart)
small vector operations, and the next step is issued before the host reads this one's
coefficients.
"""

import math
import time as _time

import numpy as _np


class IterationCounter(object):
    """Iteration counter and residual history (iterative_solvers.py:10-41)."""

    def __init__(self, store_residuals, iteration_is_cg=False, operator=None, rhs=None):
        self._count = 0
        self._store_residuals = store_residuals
        self._residuals = []
        self._iteration_is_cg = iteration_is_cg
        self._operator = operator
        self._rhs = rhs

    def __call__(self, x):
        from bempp_amd.api import log

        self._count += 1
        if self._store_residuals:
            if self._iteration_is_cg:
                res = self._rhs - self._operator(x)
            else:
                res = x
            self._residuals.append(float(_np.linalg.norm(res)))
            log(f"GMRES Iteration {self._count} with residual {self._residuals[-1]}")
        else:
            log(f"GMRES Iteration {self._count}")

    @property
    def count(self):
        return self._count

    @property
    def residuals(self):
        return self._residuals


def _lartg(f, g):
    """Givens rotation (c real, s) with [c, s; -conj(s), c] [f; g] = [r; 0] (LAPACK zlartg), on
    Python complex numbers (the per-iteration host work of GMRES)."""
    f, g = complex(f), complex(g)
    af, ag = abs(f), abs(g)
    if ag == 0.0:
        return 1.0, 0j, f
    if af == 0.0:
        return 0.0, g.conjugate() / ag, complex(ag)
    d = math.hypot(af, ag)
    ph = f / af
    return af / d, ph * g.conjugate() / d, ph * d


def device_gmres(tensor, b, rtol=1e-5, restart=None, maxiter=None, callback=None):
    """Solve tensor x = b on the device (complex128). Returns (x, info) like scipy's gmres."""
    from bempp_amd.api.assembly.discrete_boundary_operator import device_matvec

    if tensor.shape[0] != tensor.shape[1]:
        raise ValueError("GMRES needs a square operator")
    return krylov_gmres(lambda v: device_matvec(tensor, v), b, tensor.device, rtol, restart, maxiter, callback)


def krylov_gmres(matvec, b, dev, rtol=1e-5, restart=None, maxiter=None, callback=None):
    """Restarted GMRES with the operator given as matvec (complex128 tensors on dev)."""
    import torch
    from scipy.linalg import solve_triangular

    bt = torch.as_tensor(_np.asarray(b), device=dev).to(torch.complex128)
    n = bt.shape[0]
    bnrm2 = float(torch.linalg.vector_norm(bt))
    x = torch.zeros(n, dtype=torch.complex128, device=dev)
    if bnrm2 == 0.0:
        return x, 0
    atol = rtol * bnrm2
    eps = _np.finfo(_np.float64).eps
    restart = 20 if restart is None else restart
    restart = min(restart, n)
    maxiter = n * 10 if maxiter is None else maxiter
    ptol_max_factor = 1.0
    ptol = bnrm2 * min(ptol_max_factor, atol / bnrm2)
    presid = 0.0
    V = torch.empty((restart + 1, n), dtype=torch.complex128, device=dev)
    # On the GPU the CGS2 step is one library call that also writes the next basis vector, and
    # step col + 1 is issued before the host reads the coefficients of step col (alternating
    # device / pinned host buffers), so the host's rotations overlap the next product. A step
    # issued past the end of the inner loop is discarded. On the CPU the same arithmetic runs in
    # torch, one step at a time.
    hip = dev.type == "cuda"
    if hip:
        from bempp_amd.core import hip_backend as hb

        nbuf = 2 * restart + 4
        gs_out = [torch.empty(nbuf, dtype=torch.complex128, device=dev) for _ in range(2)]
        gs_host = [torch.empty(nbuf, dtype=torch.complex128, pin_memory=True) for _ in range(2)]
        gs_event = [torch.cuda.Event() for _ in range(2)]

        def issue(c):
            wv = matvec(V[c]).to(torch.complex128).contiguous()
            b = c % 2
            hb.gmres_cgs2(V, c + 1, wv, gs_out[b], V[c + 1], eps)
            gs_host[b][: 2 * c + 4].copy_(gs_out[b][: 2 * c + 4], non_blocking=True)
            gs_event[b].record()

    h = _np.zeros((restart, restart + 1), dtype=_np.complex128)
    gc, gs = [0.0] * restart, [0j] * restart  # Givens rotations (c real, s complex)
    inner_iter = 0
    rnorm = bnrm2
    r = bt.clone()
    for iteration in range(maxiter):
        if iteration == 0 and float(torch.linalg.vector_norm(r)) < atol:
            return x, 0
        tmp = float(torch.linalg.vector_norm(r))
        V[0] = r / tmp
        S = [0j] * (restart + 1)
        S[0] = complex(tmp)
        breakdown = False
        col = 0
        if hip:
            issue(0)
        for col in range(restart):
            nv = col + 1
            if hip:
                if col + 1 < restart:
                    issue(col + 1)
                gs_event[col % 2].synchronize()
                g = gs_host[col % 2][: 2 * nv + 2].numpy().copy()
                h[col, :nv] = g[:nv] + g[nv + 1 : 2 * nv + 1]
                h0v, h1v = math.sqrt(g[nv].real), math.sqrt(g[2 * nv + 1].real)
            else:
                w = matvec(V[col])
                h0 = torch.linalg.vector_norm(w)
                Vk = V[:nv]
                c1 = torch.mv(Vk.conj(), w)
                w = w - torch.mv(Vk.T, c1)
                c2 = torch.mv(Vk.conj(), w)
                w = w - torch.mv(Vk.T, c2)
                h1 = torch.linalg.vector_norm(w)
                hv = torch.cat([c1 + c2, h1.to(torch.complex128)[None], h0.to(torch.complex128)[None]]).cpu().numpy()
                h[col, :nv] = hv[:nv]
                h1v, h0v = float(hv[nv].real), float(hv[nv + 1].real)
            h[col, col + 1] = h1v
            breakdown = h1v <= eps * h0v
            if breakdown:
                h[col, col + 1] = 0.0
            if not hip:  # the device step wrote V[col + 1] itself, by the same test
                V[col + 1] = w if breakdown else w / h1v
            row = h[col, : col + 2].tolist()
            for i in range(col):
                c, s = gc[i], gs[i]
                n0, n1 = row[i], row[i + 1]
                row[i], row[i + 1] = c * n0 + s * n1, -s.conjugate() * n0 + c * n1
            c, s, mag = _lartg(row[col], row[col + 1])
            gc[col], gs[col] = c, s
            row[col], row[col + 1] = mag, 0.0
            h[col, : col + 2] = row
            tmp = -s.conjugate() * S[col]
            S[col], S[col + 1] = c * S[col], tmp
            presid = abs(tmp)
            inner_iter += 1
            if callback is not None:
                callback(presid / bnrm2)
            if inner_iter == maxiter:
                break
            if presid <= ptol or breakdown:
                break
        if h[col, col] == 0:
            S[col] = 0
        y = solve_triangular(h[: col + 1, : col + 1].T, S[: col + 1], lower=False)
        x = x + torch.mv(V[: col + 1].T, torch.as_tensor(y, device=dev))
        r = bt - matvec(x)
        rnorm = float(torch.linalg.vector_norm(r))
        if inner_iter == maxiter:
            return x, 0 if rnorm <= atol else maxiter
        if rnorm <= atol:
            break
        elif breakdown:
            break
        elif presid <= ptol:
            ptol_max_factor = max(eps, 0.25 * ptol_max_factor)
        else:
            ptol_max_factor = min(1.0, 1.5 * ptol_max_factor)
        ptol = presid * min(ptol_max_factor, atol / rnorm)
    return x, 0 if rnorm <= atol else maxiter


def device_cg(tensor, b, rtol=1e-5, maxiter=None, callback=None):
    """Conjugate gradients on the device (scipy's cg). Returns (x, info)."""
    from bempp_amd.api.assembly.discrete_boundary_operator import device_matvec

    return krylov_cg(lambda v: device_matvec(tensor, v), b, tensor.device, rtol, maxiter, callback)


def krylov_cg(matvec, b, dev, rtol=1e-5, maxiter=None, callback=None):
    """Conjugate gradients with the operator given as matvec (complex128 tensors on dev)."""
    import torch

    bt = torch.as_tensor(_np.asarray(b), device=dev).to(torch.complex128)
    n = bt.shape[0]
    bnrm2 = float(torch.linalg.vector_norm(bt))
    x = torch.zeros(n, dtype=torch.complex128, device=dev)
    if bnrm2 == 0.0:
        return x, 0
    atol = rtol * bnrm2
    maxiter = n * 10 if maxiter is None else maxiter
    r = bt.clone()
    rho_prev, p = None, None
    for iteration in range(maxiter):
        if float(torch.linalg.vector_norm(r)) < atol:
            return x, 0
        rho_cur = torch.vdot(r, r)
        if iteration > 0:
            p = p * (rho_cur / rho_prev) + r
        else:
            p = r.clone()
        q = matvec(p)
        alpha = rho_cur / torch.vdot(p, q)
        x = x + alpha * p
        r = r - alpha * q
        rho_prev = rho_cur
        if callback is not None:
            callback(x)
    return x, maxiter


def _to_host(x, real):
    xh = x.cpu().numpy()
    return xh.real.copy() if real else xh


def _is_real(A_op, b_vec):
    return not _np.issubdtype(A_op.dtype, _np.complexfloating) and not _np.iscomplexobj(b_vec)


def gmres(A, b, tol=1e-5, restart=None, maxiter=None, use_strong_form=False, return_residuals=False,
          return_iteration_count=False):
    """GMRES solve of a boundary operator or blocked operator (iterative_solvers.py:44-91)."""
    from bempp_amd.api.assembly.blocked_operator import BlockedOperator
    from bempp_amd.api.assembly.boundary_operator import BoundaryOperator

    if isinstance(A, BoundaryOperator):
        return _gmres_single_op_imp(A, b, tol, restart, maxiter, use_strong_form, return_residuals,
                                    return_iteration_count)
    if isinstance(A, BlockedOperator):
        return _gmres_block_op_imp(A, b, tol, restart, maxiter, use_strong_form, return_residuals,
                                   return_iteration_count)
    raise ValueError("A must be a BoundaryOperator or BlockedBoundaryOperator")


def _results(res_fun, info, callback, return_residuals, return_iteration_count):
    if return_residuals and return_iteration_count:
        return res_fun, info, callback.residuals, callback.count
    if return_residuals:
        return res_fun, info, callback.residuals
    if return_iteration_count:
        return res_fun, info, callback.count
    return res_fun, info


def _no_strong_form(use_strong_form):
    if use_strong_form:
        raise NotImplementedError("Strong forms (inverse mass matrix applied) are not provided; use the weak form.")


def _gmres_single_op_imp(A, b, tol, restart, maxiter, use_strong_form, return_residuals, return_iteration_count):
    from bempp_amd.api import log
    from bempp_amd.api.assembly.grid_function import GridFunction

    if not isinstance(b, GridFunction):
        raise ValueError("b must be of type GridFunction")
    _no_strong_form(use_strong_form)
    A_op = A.weak_form()
    b_vec = b.projections(A.dual_to_range)
    callback = IterationCounter(return_residuals)
    log("Starting GMRES iteration")
    start = _time.time()
    x, info = device_gmres(A_op.device_tensor(), b_vec, tol, restart, maxiter, callback)
    log("GMRES finished in %i iterations and took %.2E sec." % (callback.count, _time.time() - start))
    res_fun = GridFunction(A.domain, coefficients=_to_host(x, _is_real(A_op, b_vec)))
    return _results(res_fun, info, callback, return_residuals, return_iteration_count)


def _gmres_block_op_imp(A, b, tol, restart, maxiter, use_strong_form, return_residuals, return_iteration_count):
    from bempp_amd.api import log
    from bempp_amd.api.assembly.blocked_operator import (grid_function_list_from_coefficients,
                                                          projections_from_grid_functions_list)

    _no_strong_form(use_strong_form)
    A_op = A.weak_form()
    b_vec = projections_from_grid_functions_list(b, A.dual_to_range_spaces)
    callback = IterationCounter(return_residuals)
    log("Starting GMRES iteration")
    start = _time.time()
    x, info = device_gmres(A_op.device_tensor(), b_vec, tol, restart, maxiter, callback)
    log("GMRES finished in %i iterations and took %.2E sec." % (callback.count, _time.time() - start))
    res_fun = grid_function_list_from_coefficients(_to_host(x, _is_real(A_op, b_vec)), A.domain_spaces)
    return _results(res_fun, info, callback, return_residuals, return_iteration_count)


def cg(A, b, tol=1e-5, maxiter=None, use_strong_form=False, return_residuals=False, return_iteration_count=False):
    """CG solve of a (Hermitian positive definite) boundary operator (iterative_solvers.py:94-157)."""
    from bempp_amd.api import log
    from bempp_amd.api.assembly.boundary_operator import BoundaryOperator
    from bempp_amd.api.assembly.discrete_boundary_operator import device_matvec
    from bempp_amd.api.assembly.grid_function import GridFunction

    if not isinstance(A, BoundaryOperator):
        raise ValueError("A must be of type BoundaryOperator")
    if not isinstance(b, GridFunction):
        raise ValueError("b must be of type GridFunction")
    _no_strong_form(use_strong_form)
    A_op = A.weak_form()
    b_vec = b.projections(A.dual_to_range)
    tensor = A_op.device_tensor()

    def apply(x):
        return device_matvec(tensor, x).cpu().numpy()

    callback = IterationCounter(return_residuals, True, apply, _np.asarray(b_vec))
    log("Starting CG iteration")
    start = _time.time()
    x, info = device_cg(tensor, b_vec, tol, maxiter, callback)
    log("CG finished in %i iterations and took %.2E sec." % (callback.count, _time.time() - start))
    res_fun = GridFunction(A.domain, coefficients=_to_host(x, _is_real(A_op, b_vec)))
    return _results(res_fun, info, callback, return_residuals, return_iteration_count)
