"""Material identification on the GPU step.

``StepRun`` is a torch.autograd.Function whose forward is a recorded run of
the HIP step (``Solver.record``, mcm_gpu_adj_forward) and whose backward is
the device reverse sweep (``Recording.backward``, mcm_gpu_adj_backward).
torch only carries the parameters and the loss; every step, forward and
reverse, runs in libmcm_gpu.so.

Parameters of a model-10 material (mcm_f3dm10.f90 with eqos4): G, sigma_y
and the plastic modulus Eh (strinput(1:3)) and S1 (eosinput(2)).  The
Taylor-bar decks of mcmtop_amd/deck.py derive them from E, nu, sigma_y and
E_tan as G = E / (2 (1 + nu)), Eh = E E_tan / (E - E_tan); ``physical_to_model``
applies that map with torch so that gradients come out with respect to
(E, sigma_y, E_tan, S1).

The identification loss follows the Taylor-impact problem: squared relative
errors of the final bar length and mushroom radius, plus the sampled
surface particles' displacement histories, against a target made by the
same solver with hidden parameters.
"""
from __future__ import annotations

import numpy as np
import torch

from . import gpu


def physical_to_model(theta: torch.Tensor, nu: float) -> torch.Tensor:
    """(E, sigma_y, E_tan, S1) -> (G, sigma_y, Eh, S1)."""
    E, sy, Et, s1 = theta[0], theta[1], theta[2], theta[3]
    G = E / (2.0 * (1.0 + nu))
    Eh = E * Et / (E - Et)
    return torch.stack([G, sy, Eh, s1])


class StepRun(torch.autograd.Function):
    """(params (nummat, 4), v0 (np, 3)) -> (final positions (np, 3),
    sampled positions after every step (nsteps, nsample, 3))."""

    @staticmethod
    def forward(ctx, params, v0, solver: gpu.Solver, nsteps: int, ckpt_every: int, sample):
        p = params.detach().cpu().numpy().reshape(-1, gpu.N_PARAMS)
        for m in range(p.shape[0]):
            solver.set_material_params(m, p[m])
        solver.set_initial_velocity(v0.detach().cpu().numpy())
        solver.initial()
        need = any(ctx.needs_input_grad[:2])
        # without a gradient to take, one checkpoint and nothing kept
        rec = solver.record(nsteps, ckpt_every if need else nsteps, sample)
        x = torch.from_numpy(solver.field(gpu.F_X))
        hist = torch.from_numpy(rec.hist.copy())
        if need:
            ctx.rec = rec
        else:
            rec.close()
        return x, hist

    @staticmethod
    def backward(ctx, gx, ghist):
        rec = ctx.rec
        xb = gx.detach().cpu().numpy() if gx is not None else None
        hb = ghist.detach().cpu().numpy() if ghist is not None else None
        v0b, pb = rec.backward(xb, None, hb)
        rec.close()
        return torch.from_numpy(pb), torch.from_numpy(v0b), None, None, None, None


def bar_measures(x: torch.Tensor):
    """Final bar length (z extent) and mushroom radius (largest distance
    from the bar axis)."""
    length = x[:, 2].max() - x[:, 2].min()
    radius = torch.sqrt(x[:, 0] * x[:, 0] + x[:, 1] * x[:, 1]).max()
    return length, radius


def identification_loss(x: torch.Tensor, hist: torch.Tensor, x0s: torch.Tensor, target: dict) -> torch.Tensor:
    length, radius = bar_measures(x)
    disp = hist - x0s[None, :, :]
    return (((length - target["length"]) / target["length"]) ** 2 +
            ((radius - target["radius"]) / target["radius"]) ** 2 +
            ((disp - target["disp"]) ** 2).sum() / (target["disp"] ** 2).sum())


def surface_sample(x0: np.ndarray, radius: float, dx: float, count: int, seed: int) -> np.ndarray:
    """count distinct particles of the lateral surface (within 0.75 dx of the
    bar radius), chosen with numpy default_rng(seed)."""
    r = np.sqrt(x0[:, 0] ** 2 + x0[:, 1] ** 2)
    cand = np.nonzero(r > radius - 0.75 * dx)[0]
    rng = np.random.default_rng(seed)
    return np.sort(rng.choice(cand, size=min(count, len(cand)), replace=False)).astype(np.int32)


class Identification:
    """Loss of a Taylor bar run as a function of (E, sigma_y, E_tan, S1) and
    the initial velocities, differentiable through StepRun."""

    def __init__(self, deck, nsteps: int, ckpt_every: int, sample, nu: float, target: dict | None = None):
        self.solver = gpu.Solver(deck)
        self.nsteps = int(nsteps)
        self.ckpt_every = int(ckpt_every)
        self.sample = np.asarray(sample, dtype=np.int32)
        self.nu = nu
        self.x0 = np.asarray(deck.x, dtype=np.float64).reshape(-1, 3)
        self.v0 = np.asarray(deck.v, dtype=np.float64).reshape(-1, 3)
        self.x0s = torch.from_numpy(self.x0[self.sample])
        self.target = target

    def run(self, theta: torch.Tensor, v0: torch.Tensor):
        params = physical_to_model(theta, self.nu).reshape(1, gpu.N_PARAMS)
        return StepRun.apply(params, v0, self.solver, self.nsteps, self.ckpt_every, self.sample)

    def make_target(self, theta_hidden) -> dict:
        """Target measures from a run with the hidden parameters."""
        with torch.no_grad():
            x, hist = self.run(torch.as_tensor(theta_hidden, dtype=torch.float64), torch.from_numpy(self.v0))
            length, radius = bar_measures(x)
        self.target = {"length": length.clone(), "radius": radius.clone(), "disp": (hist - self.x0s[None]).clone()}
        return self.target

    def loss(self, theta: torch.Tensor, v0: torch.Tensor) -> torch.Tensor:
        x, hist = self.run(theta, v0)
        return identification_loss(x, hist, self.x0s, self.target)

    def value_and_grad(self, theta, v0=None):
        """loss, dL/dtheta (4), dL/dv0 (np, 3)."""
        th = torch.as_tensor(theta, dtype=torch.float64).clone().requires_grad_(True)
        v = torch.from_numpy(self.v0 if v0 is None else np.asarray(v0, dtype=np.float64)).clone().requires_grad_(True)
        L = self.loss(th, v)
        L.backward()
        return L.item(), th.grad.numpy().copy(), v.grad.numpy().copy()

    def close(self):
        self.solver.close()


class Adam:
    """Adam steps on a parameter vector (numpy)."""

    def __init__(self, n: int, lr: float, b1: float = 0.9, b2: float = 0.999, eps: float = 1e-12):
        self.lr, self.b1, self.b2, self.eps = lr, b1, b2, eps
        self.m = np.zeros(n)
        self.v = np.zeros(n)
        self.t = 0

    def step(self, x: np.ndarray, g: np.ndarray) -> np.ndarray:
        self.t += 1
        self.m = self.b1 * self.m + (1 - self.b1) * g
        self.v = self.b2 * self.v + (1 - self.b2) * g * g
        mh = self.m / (1 - self.b1 ** self.t)
        vh = self.v / (1 - self.b2 ** self.t)
        return x - self.lr * mh / (np.sqrt(vh) + self.eps)


def fit(ident: Identification, theta0, iterations: int, lr: float = 0.05, fitted=(True, True, False, True)):
    """Gradient descent (Adam on the logarithms) of the identification loss
    over the fitted parameters of (E, sigma_y, E_tan, S1), from theta0.
    Returns {"loss": [...], "theta": [...]} per iteration (the loss and
    parameters at which each gradient was taken) and the final parameters."""
    mask = np.asarray(fitted, dtype=bool)
    theta = np.asarray(theta0, dtype=np.float64).copy()
    opt = Adam(int(mask.sum()), lr)
    hist = {"loss": [], "theta": []}
    for _ in range(int(iterations)):
        L, g, _ = ident.value_and_grad(theta)
        hist["loss"].append(L)
        hist["theta"].append(theta.copy())
        u = np.log(theta[mask])
        u = opt.step(u, g[mask] * theta[mask])   # dL/dlog(theta) = theta dL/dtheta
        theta[mask] = np.exp(u)
    hist["final"] = theta
    return hist


def fit_many(idents, theta0s, iterations: int, lr: float = 0.05, threads: int = 8, fitted=(True, True, False, True),
             release: bool = False):
    """fit() for several independent instances at once: one host thread per
    instance, each run on its own HIP stream, so the GPU overlaps the
    instances' kernels (the C ABI releases the interpreter during its calls).
    release: free each instance's recording buffers when its fit ends, so
    that device memory holds those of at most ``threads`` instances."""
    from concurrent.futures import ThreadPoolExecutor

    def one(idn, th):
        out = fit(idn, th, iterations, lr, fitted)
        if release:
            idn.solver.release_recordings()
        return out

    with ThreadPoolExecutor(max_workers=max(1, int(threads))) as ex:
        futs = [ex.submit(one, idn, th) for idn, th in zip(idents, theta0s)]
        return [f.result() for f in futs]
