"""CPRSolver restatement (test infrastructure / CPU baseline only).

pymrst/modules/ad_core/solvers/linear_solvers.py:110-545 without pyamg (not
installed; the reference then uses its spilu pressure stage, :288-290):
quasi-IMPES weights 1/b_i at the mean pressure (:155-190), A_p = W A P
(:469-473), Ruiz equilibration with three sweeps (:205-224), spilu(1e-4, 10,
NATURAL) on the equilibrated A_p, cell block-Jacobi local stage on the
equilibrated cell-major matrix with an LU of the well tail (:315-363,
:382-392), and scipy gmres(rtol, restart 50, maxiter 200) with
precondition(r) = P solve_p(W r) + local(r - A P solve_p(W r)) (:510-523).
"""
import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla


def equilibrate(a, sweeps=3):
    a_s = a.tocsr().copy()
    dr = np.ones(a.shape[0])
    dc = np.ones(a.shape[1])
    for _ in range(sweeps):
        r = np.sqrt(abs(a_s).max(axis=1).toarray().ravel())
        c = np.sqrt(abs(a_s).max(axis=0).toarray().ravel())
        r[r == 0.0] = 1.0
        c[c == 0.0] = 1.0
        dr /= r
        dc /= c
        a_s = sp.diags_array(1.0 / r) @ a_s @ sp.diags_array(1.0 / c)
    return a_s, dr, dc


class CPRSolverCPU:
    def __init__(self, rtol=1e-8, maxiter=200, restart=50, use_pressure_stage=True):
        self.rtol, self.maxiter, self.restart = rtol, maxiter, restart
        self.use_pressure_stage = use_pressure_stage
        self.last_iterations = 0

    def _weighting(self, model, state, n):
        nc, nph = model.nc, model.n_phases
        p_mean = float(np.mean(state["pressure"]))
        w = np.array([1.0 / float(np.asarray(model.fluid.b(i, np.array([p_mean]))).ravel()[0])
                      for i in range(nph)])
        rows = np.tile(np.arange(nc), nph)
        cols = np.concatenate([i * nc + np.arange(nc) for i in range(nph)])
        return sp.csr_array((np.repeat(w, nc), (rows, cols)), shape=(nc, n))

    @staticmethod
    def _block_jacobi(a_i, nc, nph):
        idx = np.arange(nc) * nph
        cells = a_i[:nc * nph, :nc * nph].tocsr()
        blocks = np.empty((nc, nph, nph))
        for i in range(nph):
            for j in range(nph):
                blocks[:, i, j] = cells[idx + i, idx + j]
        inv = np.linalg.inv(blocks)
        tail = a_i[nc * nph:, nc * nph:].tocsc()
        tail_lu = None
        if tail.shape[0]:
            try:
                tail_lu = spla.splu(tail)
            except RuntimeError:
                tail_lu = spla.splu(tail + np.sqrt(np.finfo(float).eps) * sp.eye(tail.shape[0], format="csc"))

        def solve(z):
            out = np.empty_like(z)
            out[:nc * nph] = np.einsum("cij,cj->ci", inv, z[:nc * nph].reshape(nc, nph)).ravel()
            if tail_lu is not None:
                out[nc * nph:] = tail_lu.solve(z[nc * nph:])
            return out

        return solve

    def solve(self, jac, rhs, model, state):
        a = jac.tocsr()
        n = a.shape[0]
        nc, nph = model.nc, model.n_phases
        w = self._weighting(model, state, n)
        prolong = sp.csr_array((np.ones(nc), (np.arange(nc), np.arange(nc))), shape=(n, nc))
        a_p = (w @ a @ prolong).tocsr()
        a_p_eq, dr_p, dc_p = equilibrate(a_p)
        lu = spla.spilu(sp.csc_matrix(a_p_eq), drop_tol=1e-4, fill_factor=10.0, permc_spec="NATURAL")
        solve_p = lambda z: dc_p * lu.solve(dr_p * z)
        a_s, dr, dc = equilibrate(a)
        cells = (np.arange(nph) * nc + np.arange(nc)[:, None]).ravel()
        perm = np.concatenate([cells, np.arange(nc * nph, n)])
        inv_perm = np.argsort(perm)
        a_i = a_s[perm][:, perm]
        local_solve = self._block_jacobi(a_i, nc, nph)
        local = lambda z: dc * local_solve((dr * z)[perm])[inv_perm]

        def precondition(r):
            if not self.use_pressure_stage:    # the local stage alone (:487-489)
                return local(r)
            x = prolong @ solve_p(w @ r)
            return x + local(r - a @ x)

        count = [0]
        x, info = spla.gmres(a, rhs, M=spla.LinearOperator((n, n), matvec=precondition), rtol=self.rtol,
                             restart=self.restart, maxiter=self.maxiter,
                             callback=lambda _: count.__setitem__(0, count[0] + 1), callback_type="pr_norm")
        self.last_iterations = count[0]
        if info != 0:
            raise RuntimeError(f"CPR did not converge (gmres info={info})")
        return x
