"""Pseudo-particle self-consistency driver with the reference solver's interface
(python/triqs_xca/block_sparse_solver.py: BlockSparseSolver.solve, eval_pseudo_particle_self_energy,
solve_dyson, normalize_pseudo_particle_gf, partition_function_from_ppgf). Sigma evaluation, the Dyson
solve, normalization and mixing all run on the GPU; this class only holds the state."""
import numpy as np

from .evaluator import DiagramEvaluator


class BlockSparseSolver:
    def __init__(self, problem, device=0):
        self.problem = problem
        self.beta = problem.beta
        self.G0 = problem.G0
        self.eta0 = problem.eta0
        self.G = [g.copy() for g in problem.G0]
        self.eta = 0.0  # shift on top of the atomic G0 (block_sparse_solver.py:124)
        self.d = DiagramEvaluator(problem.poles, problem.coeffs, problem.grid, problem.ad, device=device)
        self.d.dyson_init(self.G0)

    def eval_pseudo_particle_self_energy(self, G, max_order):
        return self.d.pseudo_particle_self_energy(G, max_order)

    def solve_dyson(self, Sigma, eta):
        return self.d.dyson_solve(Sigma, eta)

    def solve(self, max_order, tol=1e-4, maxiter=10, mix=1.0):
        G, eta, it, diff = self.d.ppsc_solve(max_order, tol=tol, maxiter=maxiter, mix=mix, eta0=self.eta)
        self.G, self.eta, self.n_ppsc_iter, self.diff_G = G, eta, it, diff
        return self

    def single_particle_greens_function(self, max_order):
        """g(tau) from the current G (block_sparse_solver.py:653-665): (r, n, n) at the DLR nodes."""
        self.spgf = self.d.single_particle_greens_function(self.G, max_order)
        return self.spgf

    def partition_function(self):
        """Z = -Tr G(beta) exp(-beta eta)."""
        from .dlr import kernel
        grid = self.problem.grid
        kb = kernel(np.ones(1), grid.rf[None, :])[0]
        tr = sum(np.trace(np.tensordot(kb, grid.vals2coefs(g), axes=(0, 0))) for g in self.G)
        return -tr * np.exp(-self.beta * self.eta)
