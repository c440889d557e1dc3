"""Differentiable periodic box: the PIC step behind a torch.autograd.Function.

The forward step and its adjoint are HIP kernels (librunko_amd.so,
rk_diff_step / rk_diff_step_adjoint); this module holds the host side:

* ``DiffBox`` owns the library handle of one periodic box (cells, species,
  cfl, filter passes, float32 or float64) and runs one step or one adjoint
  step on torch tensors;
* ``rollout(box, E, B, x, u, steps, checkpoint_every)`` advances a state by
  ``steps`` steps as one autograd node.  The forward pass keeps the state
  every ``checkpoint_every`` steps; the backward pass recomputes each segment
  from its checkpoint, keeps the segment's states, and applies the adjoint
  steps in reverse (memory: checkpoints + one segment of states);
* the configuration of the differentiation workload: a periodic pair plasma
  whose initial magnetic field is a sum of seeded Fourier modes with
  |k| <= 3 and learnable amplitudes, species counter-drifting along x with
  a learnable drift 4-velocity, and the magnetic energy in the |k| band
  [4, 8] after the rollout as the loss.

State layout: E, B are [3, n0, n1, n2] tensors (Yee-staggered as the tile
lattice: E_x at (i+1/2, j, k), B_x at (i, j+1/2, k+1/2)); x, u are [3, Np]
with the species in consecutive ranges.
"""

import ctypes
import math

import numpy as np
import torch

from . import _lib as L
from .world import _stream


class DiffBox:
    def __init__(self, n, charges, masses, counts, cfl=0.45, filter_passes=3, dtype=torch.float32,
                 device="cuda"):
        if dtype not in (torch.float32, torch.float64):
            raise ValueError("DiffBox: dtype must be float32 or float64")
        if not (len(charges) == len(masses) == len(counts)) or len(charges) > L.RK_MAX_SPECIES:
            raise ValueError("DiffBox: charges, masses and counts must have one entry per species (at most 4)")
        self.n = tuple(int(v) for v in n)
        self.charges = [float(q) for q in charges]
        self.masses = [float(m) for m in masses]
        self.counts = [int(c) for c in counts]
        self.np = sum(self.counts)
        self.cfl = float(cfl)
        self.filter_passes = int(filter_passes)
        self.dtype = dtype
        dev = torch.device(device)
        if dev.type == "cuda" and dev.index is None:
            dev = torch.device("cuda", torch.cuda.current_device())
        self.dev = dev
        d = L.DiffDesc()
        for a in range(3):
            d.n[a] = self.n[a]
        d.n_species = len(charges)
        for s in range(len(charges)):
            d.count[s] = self.counts[s]
            d.charge[s] = self.charges[s]
            d.mass[s] = self.masses[s]
        d.cfl = self.cfl
        d.filter_passes = self.filter_passes
        d.precision = 4 if dtype == torch.float32 else 8
        h = ctypes.c_void_p()
        L.check(L.lib().rk_diff_create(ctypes.byref(h), ctypes.byref(d)))
        self.h = h
        self._fshape = (3,) + self.n

    def __del__(self):
        h = getattr(self, "h", None)
        if h is not None and h.value:
            try:
                L.lib().rk_diff_destroy(h)
            except Exception:  # interpreter shutdown: the library is already gone
                pass
            self.h = None

    # ------------------------------------------------------------------ tensors
    def empty_state(self):
        f = self._fshape
        return (torch.empty(f, dtype=self.dtype, device=self.dev), torch.empty(f, dtype=self.dtype, device=self.dev),
                torch.empty((3, self.np), dtype=self.dtype, device=self.dev),
                torch.empty((3, self.np), dtype=self.dtype, device=self.dev))

    def _check(self, state):
        E, B, x, u = state
        for t, shape in ((E, self._fshape), (B, self._fshape), (x, (3, self.np)), (u, (3, self.np))):
            if tuple(t.shape) != shape or t.dtype != self.dtype or t.device != self.dev or not t.is_contiguous():
                raise ValueError(f"DiffBox: expected a contiguous {self.dtype} tensor of shape {shape} on "
                                 f"{self.dev}, got {tuple(t.shape)} {t.dtype} {t.device}")

    @staticmethod
    def _struct(state):
        s = L.DiffState()
        for name, t in zip(("E", "B", "x", "u"), state):
            arr = getattr(s, name)
            for c in range(3):
                arr[c] = t[c].data_ptr()
        return s

    # ------------------------------------------------------------------ steps
    def step(self, state, out=None):
        """One step; returns the new state (in `out` if given; out may be state)."""
        self._check(state)
        if out is None:
            out = self.empty_state()
        self._check(out)
        a, b = self._struct(state), self._struct(out)
        L.check(L.lib().rk_diff_step(self.h, ctypes.byref(a), ctypes.byref(b), _stream()))
        return out

    def step_adjoint(self, state, adj_out, adj_in=None):
        """adj_in = J(state)^T adj_out for one step (adj_in may be adj_out)."""
        self._check(state)
        self._check(adj_out)
        if adj_in is None:
            adj_in = self.empty_state()
        self._check(adj_in)
        a, b, c = self._struct(state), self._struct(adj_out), self._struct(adj_in)
        L.check(L.lib().rk_diff_step_adjoint(self.h, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c), _stream()))
        return adj_in

    def error_count(self):
        n = ctypes.c_int64(0)
        L.check(L.lib().rk_diff_error_count(self.h, ctypes.byref(n), _stream()))
        return n.value


class DiffEnsemble(DiffBox):
    """Independent periodic boxes stepped together: every kernel of a step
    covers all boxes in one launch (rk_diff_create_ensemble).  boxes: a
    sequence of dicts with n, charges, masses, counts; cfl, filter passes and
    dtype are common.  States hold the boxes one after another: E, B are
    [3, total cells] (box i's cells from c0[i], [n0][n1][n2] inside), x, u
    are [3, total particles] (box i's from p0[i], its species consecutive);
    field_view / particle_view give one box's part."""

    def __init__(self, boxes, cfl=0.45, filter_passes=3, dtype=torch.float32, device="cuda"):
        if dtype not in (torch.float32, torch.float64):
            raise ValueError("DiffEnsemble: dtype must be float32 or float64")
        boxes = list(boxes)
        if not boxes:
            raise ValueError("DiffEnsemble: at least one box")
        self.ns, self.cells, self.nps, self.c0, self.p0 = [], [], [], [], []
        descs = (L.DiffDesc * len(boxes))()
        c0 = p0 = 0
        for i, bx in enumerate(boxes):
            n = tuple(int(v) for v in bx["n"])
            charges, masses, counts = list(bx["charges"]), list(bx["masses"]), [int(c) for c in bx["counts"]]
            if not (len(charges) == len(masses) == len(counts)) or len(charges) > L.RK_MAX_SPECIES:
                raise ValueError("DiffEnsemble: charges, masses and counts must have one entry per species (at most 4)")
            d = descs[i]
            for a in range(3):
                d.n[a] = n[a]
            d.n_species = len(charges)
            for sp in range(len(charges)):
                d.count[sp] = counts[sp]
                d.charge[sp] = float(charges[sp])
                d.mass[sp] = float(masses[sp])
            d.cfl = float(cfl)
            d.filter_passes = int(filter_passes)
            d.precision = 4 if dtype == torch.float32 else 8
            self.ns.append(n)
            self.cells.append(int(np.prod(n)))
            self.nps.append(sum(counts))
            self.c0.append(c0)
            self.p0.append(p0)
            c0 += self.cells[-1]
            p0 += self.nps[-1]
        self.ncell, self.np = c0, p0
        self.cfl = float(cfl)
        self.filter_passes = int(filter_passes)
        self.dtype = dtype
        dev = torch.device(device)
        if dev.type == "cuda" and dev.index is None:
            dev = torch.device("cuda", torch.cuda.current_device())
        self.dev = dev
        h = ctypes.c_void_p()
        L.check(L.lib().rk_diff_create_ensemble(ctypes.byref(h), descs, len(boxes)))
        self.h = h
        self._fshape = (3, self.ncell)

    @property
    def n_boxes(self):
        return len(self.ns)

    def field_view(self, F, i):
        return F[:, self.c0[i]:self.c0[i] + self.cells[i]].view((3,) + self.ns[i])

    def particle_view(self, x, i):
        return x[:, self.p0[i]:self.p0[i] + self.nps[i]]


def checkpoint_levels(every):
    """Normalised checkpoint spacing: an int K (a checkpoint every K steps,
    each segment's states recomputed and kept whole in the backward pass) or
    a decreasing tuple (K1, K2, ...): checkpoints every K1 steps in the
    forward pass, and each K1 segment recomputed with checkpoints every K2
    steps, and so on, the last level kept whole.  Peak memory is about
    steps/K1 + K1/K2 + ... + K_last states; every level below the first
    adds one forward pass of recomputation."""
    levels = (int(every),) if np.isscalar(every) else tuple(int(k) for k in every)
    if not levels or any(k < 1 for k in levels) or any(a < b for a, b in zip(levels, levels[1:])):
        raise ValueError("rollout: checkpoint spacing must be a positive int or a decreasing tuple of them")
    return levels


def checkpoint_states(steps, every):
    """States held at the peak of a rollout's backward pass (the checkpoints
    of every level plus the last level's whole segment)."""
    levels = checkpoint_levels(every)
    n, out = int(steps), 0
    for k in levels:
        out += -(-n // k)
        n = min(k, n)
    return out + n


def choose_checkpoints(steps, state_bytes, budget_bytes, reserve_states=4, max_levels=4):
    """The checkpoint spacing with the least recomputation whose peak states
    (plus `reserve_states` working states) fit budget_bytes: flat when it
    fits, else the fewest levels that fit (each level below the first costs
    one more forward pass), and among those the spacing holding the fewest
    states."""
    steps = int(steps)
    fit = int(budget_bytes // max(int(state_bytes), 1)) - reserve_states
    memo = {}

    def best(n, depth):
        # (states, spacing) for a segment of n steps with `depth` levels
        if depth == 0 or n <= 1:
            return n, ()
        key = (n, depth)
        if key not in memo:
            out = (n, ())  # kept whole
            for k in range(1, n):
                c, sub = best(k, depth - 1)
                c += -(-n // k)
                if c < out[0]:
                    out = (c, (k,) + sub)
            memo[key] = out
        return memo[key]

    for depth in range(1, max_levels + 1):
        c, spacing = best(steps, depth)
        if c <= fit:
            return spacing[0] if len(spacing) == 1 else spacing
    raise RuntimeError(f"rollout: {budget_bytes / 2**30:.1f} GiB hold {fit} states of {state_bytes / 2**30:.2f} "
                       f"GiB beside the working ones; no spacing of up to {max_levels} levels fits")


def _adjoint_segment(box, s0, n, levels, lam):
    """lam holds the adjoint of the state n steps after s0; on return it
    holds the adjoint of s0.  The forward states are recomputed from s0:
    kept whole when no level is left, else with checkpoints every levels[0]
    steps, each sub-segment handled by the next level."""
    if not levels or levels[0] >= n:
        states = [s0]
        for _ in range(n - 1):
            states.append(box.step(states[-1]))
        for st in reversed(states):
            box.step_adjoint(st, lam, lam)
        return
    k = levels[0]
    ckpts = [s0]
    s = s0
    for t in range(k, n, k):
        for i in range(k):
            s = box.step(s)
        ckpts.append(s)
    del s
    for i in reversed(range(len(ckpts))):
        _adjoint_segment(box, ckpts[i], min(k, n - i * k), levels[1:], lam)
        ckpts[i] = None


class _Rollout(torch.autograd.Function):
    @staticmethod
    def forward(ctx, box, steps, every, E, B, x, u):
        levels = checkpoint_levels(every)
        s = tuple(t.detach().contiguous().clone() for t in (E, B, x, u))
        spare = box.empty_state()
        ckpts = []
        for t in range(steps):
            if t % levels[0] == 0:
                ckpts.append(tuple(v.clone() for v in s))
            box.step(s, spare)
            s, spare = spare, s
        ctx.box, ctx.steps, ctx.levels, ctx.ckpts = box, steps, levels, ckpts
        return s

    @staticmethod
    def backward(ctx, gE, gB, gx, gu):
        box, steps, levels = ctx.box, ctx.steps, ctx.levels
        ref = box.empty_state()
        lam = tuple(torch.zeros_like(r) if g is None else g.detach().contiguous().clone()
                    for g, r in zip((gE, gB, gx, gu), ref))
        del ref
        k = levels[0]
        for seg in reversed(range(len(ctx.ckpts))):
            t0 = seg * k
            n = min(t0 + k, steps) - t0
            s0 = ctx.ckpts[seg]
            ctx.ckpts[seg] = None
            _adjoint_segment(box, s0, n, levels[1:], lam)
            del s0
        return (None, None, None) + lam


def rollout(box, E, B, x, u, steps, checkpoint_every=20):
    """State after `steps` steps, differentiable w.r.t. E, B, x, u.
    checkpoint_every: an int, or a decreasing tuple for nested checkpoints
    (checkpoint_levels)."""
    if steps < 1:
        raise ValueError("rollout: steps and checkpoint_every must be positive")
    checkpoint_levels(checkpoint_every)
    return _Rollout.apply(box, int(steps), checkpoint_every, E, B, x, u)


# ---------------------------------------------------------------------- the workload
def mode_table(n_modes=12, kmax=3, seed=7):
    """Seeded Fourier modes of the initial magnetic field: integer wave vectors
    0 < |k| <= kmax (k and -k not both), unit polarisations perpendicular to
    k, phases in [0, 2 pi)."""
    rng = np.random.default_rng(seed)
    ks, es, ph = [], [], []
    while len(ks) < n_modes:
        k = rng.integers(-kmax, kmax + 1, 3)
        kk = float(np.sqrt((k * k).sum()))
        if kk == 0 or kk > kmax or any(np.array_equal(k, q) or np.array_equal(-k, q) for q in ks):
            continue
        r = rng.standard_normal(3)
        e = np.cross(k, r)
        e /= np.linalg.norm(e)
        ks.append(k)
        es.append(e)
        ph.append(rng.uniform(0, 2 * np.pi))
    return np.array(ks, dtype=np.float64), np.array(es), np.array(ph)


def staggered_coords(n, component, field, device):
    """Coordinates (float64, [3, n0, n1, n2]) where component `component` of
    E or B lives: E_c at +1/2 along c; B_c at +1/2 along the two other axes."""
    g = torch.meshgrid(*(torch.arange(v, dtype=torch.float64, device=device) for v in n), indexing="ij")
    r = [g[0].clone(), g[1].clone(), g[2].clone()]
    for a in range(3):
        if (field == "E" and a == component) or (field == "B" and a != component):
            r[a] += 0.5
    return r


def initial_B(amp, table, n, B0=(0.0, 0.0, 0.0), dtype=torch.float32):
    """B = B0 + sum_m amp_m e_m cos(2 pi k_m . r / n + phi_m) on the B stagger,
    differentiable w.r.t. amp (float64 inside, cast to dtype)."""
    k, e, ph = table
    dev = amp.device
    kt = torch.as_tensor(k, device=dev)
    et = torch.as_tensor(e, device=dev)
    pt = torch.as_tensor(ph, device=dev)
    out = []
    for c in range(3):
        r = staggered_coords(n, c, "B", dev)
        arg = sum(kt[:, a].view(-1, 1, 1, 1) * (2 * math.pi / n[a]) * r[a].unsqueeze(0) for a in range(3))
        arg = arg + pt.view(-1, 1, 1, 1)
        val = ((amp.to(torch.float64) * et[:, c]).view(-1, 1, 1, 1) * torch.cos(arg)).sum(0) + B0[c]
        out.append(val)
    return torch.stack(out).to(dtype)


def band_magnetic_energy(B, kmin=4.0, kmax=8.0):
    """sum over wave vectors with kmin <= |k| <= kmax of |B_k|^2 / (2 N), N the
    cell count (Parseval: the share of sum(B^2)/2 in that band)."""
    n = B.shape[1:]
    Bh = torch.fft.fftn(B.to(torch.float64), dim=(1, 2, 3))
    f = [torch.fft.fftfreq(v, d=1.0 / v, device=B.device, dtype=torch.float64) for v in n]
    kk = torch.sqrt(f[0].view(-1, 1, 1) ** 2 + f[1].view(1, -1, 1) ** 2 + f[2].view(1, 1, -1) ** 2)
    mask = ((kk >= kmin) & (kk <= kmax)).to(torch.float64)
    N = float(np.prod(n))
    return 0.5 * ((Bh.real ** 2 + Bh.imag ** 2) * mask).sum() / N


class PairPlasmaProblem:
    """Periodic pair plasma with learnable initial dB mode amplitudes and
    counter-drift: species 0 gets +u_d along x, species 1 -u_d, on top of
    isotropic Juttner-Synge momenta (theta), both species at the same
    uniformly random positions, ppc particles per cell per species, charges
    of projects/pic-turbulence/pic.py:44-53 (c/omega_p = c_omp cells)."""

    def __init__(self, n=(96, 96, 96), ppc=16, theta=0.3, c_omp=10.0, cfl=0.45, n_modes=12, kmax=3, seed=11,
                 filter_passes=3, dtype=torch.float32, device="cuda", amp_scale=None):
        from .init import plasma_parameters, juttner_momenta
        self.n = tuple(n)
        self.cfl = cfl
        self.dtype = dtype
        self.dev = torch.device(device)
        q0, q1, Bsig = plasma_parameters(cfl, c_omp, ppc, theta, 1.0)
        self.charges = [q0, q1]
        ncell = int(np.prod(self.n))
        npart = ncell * ppc
        self.box = DiffBox(self.n, self.charges, [1.0, 1.0], [npart, npart], cfl=cfl, filter_passes=filter_passes,
                           dtype=dtype, device=device)
        self.table = mode_table(n_modes, kmax, seed)
        self.amp_scale = Bsig / math.sqrt(n_modes) if amp_scale is None else amp_scale
        gen = torch.Generator(device=self.dev)
        gen.manual_seed(seed)
        cells = torch.stack(torch.meshgrid(*(torch.arange(v, device=self.dev, dtype=torch.float64) for v in self.n),
                                           indexing="ij")).reshape(3, -1)
        pos = cells.repeat_interleave(ppc, dim=1) + torch.rand((3, npart), generator=gen, device=self.dev,
                                                               dtype=torch.float64)
        # positions drawn in float32, so float32 and float64 problems of one seed start from the same particles
        self.x0 = torch.cat([pos, pos], 1).to(torch.float32).to(dtype).contiguous()
        self.u_thermal = torch.cat([juttner_momenta(npart, theta, gen, self.dev).to(torch.float64)
                                    for _ in range(2)], 1)
        self.sign = torch.cat([torch.ones(npart, device=self.dev, dtype=torch.float64),
                               -torch.ones(npart, device=self.dev, dtype=torch.float64)])

    def initial_state(self, amp, drift):
        """(E, B, x, u) at t = 0 from mode amplitudes [n_modes] (in units of
        amp_scale) and the drift 4-velocity (scalar tensors, any float type)."""
        B = initial_B(amp * self.amp_scale, self.table, self.n, dtype=self.dtype)
        E = torch.zeros((3,) + self.n, dtype=self.dtype, device=self.dev)
        ux = self.u_thermal[0] + self.sign * drift.to(torch.float64)
        u = torch.stack([ux, self.u_thermal[1], self.u_thermal[2]]).to(self.dtype)
        return E, B.contiguous(), self.x0, u.contiguous()

    def loss(self, amp, drift, steps, checkpoint_every=20):
        E, B, x, u = self.initial_state(amp, drift)
        E, B, x, u = rollout(self.box, E, B, x, u, steps, checkpoint_every)
        return band_magnetic_energy(B)


class EnsembleProblem:
    """The batched ensemble workload: n_boxes independent periodic pair
    plasmas, each with a cubic side drawn from `sides`, ppc particles per
    cell per species drawn from `ppcs`, theta log-uniform in theta_range,
    sigma log-uniform in sigma_range (charges and guide field B0 along z of
    projects/pic-turbulence/pic.py:44-68 with that box's ppc, theta, sigma),
    and n_modes seeded dB modes (|k| <= kmax) with learnable amplitudes
    [n_boxes, n_modes] in units of B0 / sqrt(n_modes).  The loss is the sum
    over boxes of each box's final magnetic energy sum(B^2)/2: the boxes are
    independent, so its gradient holds every box's own gradient."""

    def __init__(self, n_boxes=64, sides=(32, 48, 64, 96), ppcs=(4, 8, 16), theta_range=(0.05, 3.0),
                 sigma_range=(0.1, 10.0), n_modes=6, kmax=3, c_omp=10.0, cfl=0.45, seed=5, filter_passes=3,
                 dtype=torch.float32, device="cuda"):
        from .init import plasma_parameters, juttner_momenta
        rng = np.random.default_rng(seed)
        self.dtype = dtype
        self.dev = torch.device(device)
        self.n_modes = n_modes
        self.params = []
        boxes = []
        for i in range(n_boxes):
            side = int(rng.choice(sides))
            ppc = int(rng.choice(ppcs))
            theta = float(np.exp(rng.uniform(np.log(theta_range[0]), np.log(theta_range[1]))))
            sigma = float(np.exp(rng.uniform(np.log(sigma_range[0]), np.log(sigma_range[1]))))
            q0, q1, B0 = plasma_parameters(cfl, c_omp, ppc, theta, sigma)
            npart = side ** 3 * ppc
            self.params.append(dict(side=side, ppc=ppc, theta=theta, sigma=sigma, q0=q0, B0=B0,
                                    table=mode_table(n_modes, kmax, seed=int(rng.integers(1 << 30)))))
            boxes.append(dict(n=(side,) * 3, charges=[q0, q1], masses=[1.0, 1.0], counts=[npart, npart]))
        self.box = DiffEnsemble(boxes, cfl=cfl, filter_passes=filter_passes, dtype=dtype, device=device)
        gen = torch.Generator(device=self.dev)
        gen.manual_seed(seed)
        xs, us = [], []
        for p in self.params:
            side, ppc = p["side"], p["ppc"]
            npart = side ** 3 * ppc
            cells = torch.stack(torch.meshgrid(*(torch.arange(side, device=self.dev, dtype=torch.float64)
                                                 for _ in range(3)), indexing="ij")).reshape(3, -1)
            pos = cells.repeat_interleave(ppc, dim=1) + torch.rand((3, npart), generator=gen, device=self.dev,
                                                                   dtype=torch.float64)
            pos = pos.to(torch.float32)
            xs.append(torch.cat([pos, pos], 1))
            us.append(torch.cat([juttner_momenta(npart, p["theta"], gen, self.dev) for _ in range(2)], 1))
            del cells, pos
        self.x0 = torch.cat(xs, 1).to(dtype).contiguous()
        self.u0 = torch.cat(us, 1).to(dtype).contiguous()
        del xs, us

    def initial_state(self, amp):
        """(E, B, x, u) at t = 0 from the mode amplitudes [n_boxes, n_modes]."""
        Bs = []
        for i, p in enumerate(self.params):
            n = (p["side"],) * 3
            scale = p["B0"] / math.sqrt(self.n_modes)
            Bs.append(initial_B(amp[i] * scale, p["table"], n, B0=(0.0, 0.0, p["B0"]),
                                dtype=self.dtype).reshape(3, -1))
        B = torch.cat(Bs, 1).contiguous()
        E = torch.zeros_like(B)
        return E, B, self.x0, self.u0

    def box_magnetic_energies(self, B):
        """[n_boxes] float64: sum(B^2)/2 of every box."""
        return torch.stack([0.5 * (self.box.field_view(B, i).to(torch.float64) ** 2).sum()
                            for i in range(self.box.n_boxes)])

    def loss(self, amp, steps, checkpoint_every=10):
        E, B, x, u = self.initial_state(amp)
        E, B, x, u = rollout(self.box, E, B, x, u, steps, checkpoint_every)
        return self.box_magnetic_energies(B).sum()
