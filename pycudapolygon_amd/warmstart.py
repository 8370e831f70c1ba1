"""Learned warm start for FIRE minimization (host side of the PCP_MPNN_* entry points).

No counterpart in the reference (camdennis/pyCudaPolygon has no learned component).
A small message-passing network over the contact graph of a packing predicts a
displacement of every vertex; the displacement is applied once, the polygons are put
back on their shapes by SHAKE, and plain FP64 FIRE then minimizes to max|F| <= fTol
and judges the residual force.  Network, features and kernels: include/pcp.h
(PCP_MPNN_*) and csrc/pcp_mpnn.hip; everything here runs on the GPU through the C ABI.

Training data are FIRE trajectories that the method itself produces: states captured
at a few step counts along each trajectory, each labelled with the displacement (minimum
image) from that state to the packing's minimized state.  Training is mean squared error
with Adam, on the GPU, through the library's own forward and backward kernels.
"""
import ctypes
import os
import time

import numpy as np

from . import capi
from .generate import make_packing, box_size_for

NPARAM = 25218
DEG = 16
H = 64
SQUARED = capi.NORMAL | 0x200
PAIR_SQUARED = capi.NORMAL | 0x400
# weights trained by `python -m pycudapolygon_amd.train_warmstart` (its defaults)
DEFAULT_WEIGHTS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "weights", "warmstart_n1024_n16.npz")
# weights trained on packing-fraction steps of the compression protocol:
# `python -m pycudapolygon_amd.train_warmstart --mode compression --N 4096 --n 24 --replicas 16
#  --phi-steps 40 --epochs 10 --seed 2000`
COMPRESS_WEIGHTS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "weights",
                                "warmstart_compress_n4096_n24.npz")
# FIRE time-step cap for the squared lens energy (see train_warmstart)
DT_MAX = 100.0
# pcp_mpnn_guard: a packing keeps its input positions when the warm start puts one polygon's
# bounding circle inside another's, |c_A - c_B| + r_small < GUARD_KAPPA * r_large (the
# predicted displacements routinely bring centres to 0.2-0.4 of the radius sum, which FIRE
# resolves; a polygon whose circle lies inside its neighbour's is pulled all the way in)
GUARD_KAPPA = 1.0


def _np_ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def init_params(seed):
    """He-style initial parameters (float32), biases zero; layout of include/pcp.h."""
    rng = np.random.default_rng(seed)
    p = np.zeros(NPARAM, np.float32)
    p[0:256] = rng.normal(0, np.sqrt(2.0 / 4), 256)
    for l in range(3):
        o = 320 + l * 8256
        p[o:o + 8192] = rng.normal(0, 0.5 * np.sqrt(2.0 / 128), 8192)
    p[25088:25216] = rng.normal(0, np.sqrt(1.0 / 64), 128)
    return p


def packing_length_units(packings):
    """ell of each packing: its mean target edge length."""
    return np.array([float(np.mean(p["targetEdgeV"])) for p in packings])


class Network:
    """Device-resident parameters plus forward/backward work buffers for up to Vmax vertices."""

    def __init__(self, ctx, params, Vmax, train=False):
        self.ctx = ctx
        self.params = capi.DeviceArray.from_numpy(np.asarray(params, np.float32))
        self.Vmax = 0
        self.train = train
        self._grow(Vmax)
        if train:
            self.grad = capi.DeviceArray((NPARAM,), np.float32)
            self.m = capi.DeviceArray((NPARAM,), np.float32); self.m.zero()
            self.v = capi.DeviceArray((NPARAM,), np.float32); self.v.zero()
            self.step = 0

    def _grow(self, V):
        if V <= self.Vmax:
            return
        wf, gf = ctypes.c_longlong(0), ctypes.c_longlong(0)
        capi.check(capi.lib().pcp_mpnn_sizes(int(V), ctypes.byref(wf), ctypes.byref(gf)))
        self.work = capi.DeviceArray((wf.value,), np.float32)
        self.out = capi.DeviceArray((2 * V,), np.float32)
        if self.train:
            self.gwork = capi.DeviceArray((gf.value,), np.float32)
            self.dout = capi.DeviceArray((2 * V,), np.float32)
        self.Vmax = V

    def sync(self):
        """wait for the context's stream (device buffers are then safe to read back)"""
        self.ctx.call("pcp_synchronize")

    def get_params(self):
        self.sync()
        return self.params.get()

    def forward(self, V, feat_ptr, adj_ptr):
        """queue one forward pass on the context's stream (asynchronous)"""
        self._grow(V)
        self.ctx.call("pcp_mpnn_forward", int(V), self.params.ptr, feat_ptr, adj_ptr, self.work.ptr, self.out.ptr)
        return self.out

    def train_step(self, V, feat_ptr, adj_ptr, target_ptr, lr, want_loss=False):
        """forward, loss, backward and one Adam step; returns the loss (or None)."""
        self.forward(V, feat_ptr, adj_ptr)
        loss = ctypes.c_double(0.0)
        self.ctx.call("pcp_mpnn_loss", int(V), self.out.ptr, target_ptr, self.dout.ptr, self.gwork.ptr,
                      ctypes.byref(loss) if want_loss else None)
        self.ctx.call("pcp_mpnn_backward", int(V), self.params.ptr, feat_ptr, adj_ptr, self.work.ptr,
                      self.dout.ptr, self.gwork.ptr, self.grad.ptr)
        self.step += 1
        self.ctx.call("pcp_mpnn_adam", self.params.ptr, self.grad.ptr, self.m.ptr, self.v.ptr, self.step,
                      ctypes.c_float(lr), ctypes.c_float(0.9), ctypes.c_float(0.999), ctypes.c_float(1e-8))
        return loss.value if want_loss else None

    def loss(self, V, feat_ptr, adj_ptr, target_ptr):
        self._grow(V)
        if not self.train:
            raise RuntimeError("loss needs a training network")
        self.forward(V, feat_ptr, adj_ptr)
        loss = ctypes.c_double(0.0)
        self.ctx.call("pcp_mpnn_loss", int(V), self.out.ptr, target_ptr, self.dout.ptr, self.gwork.ptr,
                      ctypes.byref(loss))
        return loss.value


class GraphFeatures:
    """Contact graph + input features of a System's current evaluation (device buffers)."""

    def __init__(self, system, ell, inScale, energyPow=4):
        self.s = system
        self.adj = capi.DeviceArray((system.V * DEG,), np.int32)
        self.feat = capi.DeviceArray((system.V * 4,), np.float32)
        self.ell = capi.DeviceArray.from_numpy(np.asarray(ell, np.float64))
        self.inScale = np.asarray(inScale, np.float32)
        self.energyPow = int(energyPow)

    def compute(self, want_dropped=False):
        d = ctypes.c_int(0)
        self.s.ctx.call("pcp_mpnn_graph", self.adj.ptr, ctypes.byref(d) if want_dropped else None)
        self.s.ctx.call("pcp_mpnn_features", ctypes.byref(self.s._state), self.adj.ptr, self.ell.ptr,
                        self.energyPow, _np_ptr(self.inScale), self.feat.ptr)
        self.s.sync()
        return d.value


class WarmStart:
    """A trained network applied to a System: features at the current positions (after an
    evaluation), one forward pass, x += ell * outScale * y, then SHAKE."""

    @classmethod
    def default(cls):
        return cls(load_weights(DEFAULT_WEIGHTS))

    def __init__(self, weights):
        self.params = np.asarray(weights["params"], np.float32)
        self.inScale = np.asarray(weights["inScale"], np.float32)
        self.outScale = float(weights["outScale"])
        self.energyPow = int(weights.get("energyPow", 4))
        self.shakeIter = int(weights.get("shakeIter", 1000))

    def apply(self, system, ell, active=None):
        """Apply once to every packing of `system` whose forces were just evaluated.
        A packing in which the prediction puts one polygon's bounding circle inside
        another's keeps its input positions (pcp_mpnn_guard, GUARD_KAPPA: no edge-crossing
        energy sees a polygon inside another); their number is left in self.taken_back.  Returns the seconds spent (device work
        included)."""
        t0 = time.perf_counter()
        gf = GraphFeatures(system, ell, self.inScale, self.energyPow)
        gf.compute()
        net = Network(system.ctx, self.params, system.V)
        out = net.forward(system.V, gf.feat.ptr, gf.adj.ptr)
        prev = capi.DeviceArray((2 * system.V,), np.float64)
        capi.check(capi.lib().pcp_memcpy_d2d(prev.ptr, system.pos.ptr, 16 * system.V))
        system.ctx.call("pcp_mpnn_apply", system.pos.ptr, out.ptr, gf.ell.ptr, self.outScale,
                        active.ptr if active is not None else None)
        system.shake(self.shakeIter, 1e-15)
        flags = capi.DeviceArray((max(system.P, 1),), np.int32)
        n = ctypes.c_int(0)
        system.ctx.call("pcp_mpnn_guard", system.pos.ptr, prev.ptr, GUARD_KAPPA, flags.ptr, ctypes.byref(n))
        self.taken_back = n.value
        system.sync()
        return time.perf_counter() - t0


def save_weights(path, params, inScale, outScale, energyPow=4, **meta):
    np.savez(path, params=np.asarray(params, np.float32), inScale=np.asarray(inScale, np.float32),
             outScale=np.float64(outScale), energyPow=np.int64(energyPow),
             meta=np.array(repr(meta)))


def load_weights(path):
    with np.load(path, allow_pickle=False) as z:
        return dict(params=z["params"], inScale=z["inScale"], outScale=float(z["outScale"]),
                    energyPow=int(z["energyPow"]))


# ---- training data ---------------------------------------------------------------------
def make_c4_packings(P, seed, N=1024, n=16, phi_range=(0.82, 0.88)):
    """P packings of the learned-warm-start configuration: N polygons of n vertices,
    phi ~ U[phi_range], centres at least (R_i + R_j)/2 apart."""
    rng = np.random.default_rng(seed)
    phis = rng.uniform(phi_range[0], phi_range[1], P)
    seeds = rng.integers(0, 2 ** 31 - 1, P)
    return [make_packing(N, n, float(ph), int(sd), min_center_frac=0.5) for ph, sd in zip(phis, seeds)]


def fire_system(packings, fTol=1e-10, dtMax=0.1):
    """System of the packings (squared lens areas, rigid polygons) and its FIRE parameters
    (the reference driver's defaults except dtMax)."""
    B = [box_size_for(p["diameter"]) for p in packings]
    s = capi.System(packings, B, model=SQUARED)
    prm = s.fire_params(fTol=fTol, dtMax=dtMax)
    return s, prm


def fire_run(s, prm, maxSteps, dt0=1e-3, init=True):
    """FIRE from the current positions until every packing converges or maxSteps; returns
    (steps per packing (-1 = not converged), total steps, seconds)."""
    if init:
        s.fire_init(prm, dt0)
    conv = np.full(s.P, -1)
    t0 = time.perf_counter()
    steps = 0
    while steps < maxSteps:
        st = s.fire_step(prm)
        steps += 1
        for k in range(s.P):
            if conv[k] < 0 and st[k].converged:
                conv[k] = st[k].steps
        if (conv >= 0).all():
            break
    s.sync()
    return conv, steps, time.perf_counter() - t0


class Dataset:
    """Captured states of FIRE trajectories, on the device: per sample (one System state of
    P packings) features [V,4], adjacency [V,DEG] (ids local to the sample), positions [2V]
    and, once labelled, targets [V,2]."""

    def __init__(self):
        self.samples = []   # list of dict(feat, adj, x, target, V)

    def capture(self, s, gf):
        gf.compute()
        V = s.V
        feat = capi.DeviceArray((V * 4,), np.float32)
        adj = capi.DeviceArray((V * DEG,), np.int32)
        x = capi.DeviceArray((2 * V,), np.float64)
        capi.check(capi.lib().pcp_memcpy_d2d(feat.ptr, gf.feat.ptr, feat.nbytes))
        capi.check(capi.lib().pcp_memcpy_d2d(adj.ptr, gf.adj.ptr, adj.nbytes))
        capi.check(capi.lib().pcp_memcpy_d2d(x.ptr, s.pos.ptr, x.nbytes))
        self.samples.append(dict(feat=feat, adj=adj, x=x, target=None, V=V, sys=s, ell=gf.ell))

    def label(self, s, outScale):
        """targets of every sample of system s = displacement to its own end state (xto, when
        the sample has one) or else to s's current positions"""
        for d in self.samples:
            if d["sys"] is not s or d["target"] is not None:
                continue
            d["target"] = capi.DeviceArray((2 * d["V"],), np.float32)
            to = d["xto"].ptr if d.get("xto") is not None else s.pos.ptr
            s.ctx.call("pcp_mpnn_targets", d["x"].ptr, to, d["ell"].ptr, float(outScale), d["target"].ptr)
        s.sync()

    def set_end_state(self, s, first):
        """the end state of samples[first:] of system s is s's current positions"""
        for d in self.samples[first:]:
            if d["sys"] is s:
                d["xto"] = capi.DeviceArray((2 * d["V"],), np.float64)
                capi.check(capi.lib().pcp_memcpy_d2d(d["xto"].ptr, s.pos.ptr, d["xto"].nbytes))


def raw_feature_scales(s, ell, energyPow=4):
    """Input scales from a System's current evaluation: RMS of F/ell^(e-1) per component,
    1 for the crossing count, RMS of E_poly/(n ell^e) over the polygons with energy."""
    F = s.get_forces().reshape(-1, 2)
    lv = np.repeat(np.asarray(ell), np.diff(s.packVert))
    Fs = F / lv[:, None] ** (energyPow - 1)
    rms = lambda a: float(np.sqrt(np.mean(a.astype(np.float64) ** 2))) if len(a) and np.any(a) else 1.0
    sc = np.array([rms(Fs[:, 0]), rms(Fs[:, 1]), 1.0, 1.0], np.float32)
    gf = GraphFeatures(s, ell, sc, energyPow)
    gf.compute()
    e = np.sinh(gf.feat.get().reshape(-1, 4)[:, 3].astype(np.float64))
    sc[3] = rms(e[e > 0])
    return sc


def generate(batches, P, seed, capture_steps=(0, 100, 300, 1000), maxSteps=12000, log=print, N=1024, n=16,
             inScale=None, dtMax=0.1):
    """Run FIRE on `batches` systems of P packings (make_c4_packings) and capture states at
    capture_steps.  Returns (Dataset, runs = [(system, ell, conv, seconds)], inScale); the
    input scales come from the first system's initial evaluation unless given."""
    ds = Dataset()
    runs = []
    for b in range(batches):
        pk = make_c4_packings(P, seed + 7919 * b, N=N, n=n)
        ell = packing_length_units(pk)
        s, prm = fire_system(pk, dtMax=dtMax)
        s.fire_init(prm, 1e-3)
        if inScale is None:
            inScale = raw_feature_scales(s, ell)
        gf = GraphFeatures(s, ell, inScale)
        conv = np.full(s.P, -1)
        t0 = time.perf_counter()
        steps = 0
        caps = sorted(set(capture_steps))
        if 0 in caps:
            ds.capture(s, gf)
        while steps < maxSteps:
            st = s.fire_step(prm)
            steps += 1
            for k in range(s.P):
                if conv[k] < 0 and st[k].converged:
                    conv[k] = st[k].steps
            if steps in caps:
                ds.capture(s, gf)
            if (conv >= 0).all():
                break
        el = time.perf_counter() - t0
        runs.append((s, ell, conv, el))
        ok = conv >= 0
        log(f"[warmstart] batch {b}: {P} packings, FIRE {steps} steps in {el:.1f} s, converged "
            f"{int(ok.sum())}/{P}, median steps {int(np.median(conv[ok])) if ok.any() else -1}")
    return ds, runs, inScale


def generate_compression(replicas, N, n, seed, phi0=0.80, steps=40, dphi=1e-3, maxSteps=3000, dtMax=DT_MAX,
                         inScale=None, log=print):
    """Training data of the compression protocol: `replicas` packings minimized at phi0, then
    `steps` packing-fraction steps (targets scaled, polygons rescaled about their centroids);
    each just-rescaled state is captured (after one evaluation) and labelled with the
    minimized state that FIRE reaches from it.  Returns (Dataset, runs, inScale)."""
    from . import compress
    pk = [make_packing(N, n, phi0, seed + k, min_center_frac=0.5) for k in range(replicas)]
    phi_end = phi0 + steps * dphi
    B = [box_size_for(p["diameter"] * np.sqrt(phi_end / phi0)) for p in pk]
    s = capi.System(pk, B, model=SQUARED)
    prm = s.fire_params(fTol=1e-10, dtMax=dtMax)
    ell = packing_length_units(pk)
    s.fire_init(prm, 1e-3)
    s.fire_run(prm, maxSteps, 8, True)
    ds = Dataset()
    phi = compress.packing_fractions(s)
    t0 = time.perf_counter()
    fire_steps = []
    for k in range(steps):
        prev = phi
        phi = phi + dphi
        compress.set_phi(s, phi)
        ell = ell * np.sqrt(phi / prev)
        s.force_eval()
        if inScale is None:
            inScale = raw_feature_scales(s, ell)
        gf = GraphFeatures(s, ell, inScale)
        first = len(ds.samples)
        ds.capture(s, gf)
        s.fire_init(prm, 1e-3)
        st = s.fire_run(prm, maxSteps, 8, True)
        fire_steps.append([q.steps for q in st])
        ds.set_end_state(s, first)
    el = time.perf_counter() - t0
    conv = np.array(fire_steps[-1])
    log(f"[warmstart] compression data: {replicas} x N={N} x {n}, {steps} steps of {dphi}, "
        f"mean FIRE steps per step {np.mean(fire_steps):.0f}, {el:.1f} s")
    return ds, [(s, ell, conv, el)], inScale


def label_all(ds, runs):
    """Label every sample with the displacement to its system's final positions, scaled so
    that the targets have unit RMS per component; returns that scale (outScale)."""
    for s, _, _, _ in runs:
        ds.label(s, 1.0)
    acc, cnt = 0.0, 0
    for d in ds.samples:
        t = d["target"].get().astype(np.float64)
        acc += float((t ** 2).sum())
        cnt += t.size
    outScale = float(np.sqrt(acc / max(cnt, 1))) or 1.0
    for d in ds.samples:
        d["target"] = None
    for s, _, _, _ in runs:
        ds.label(s, outScale)
    return outScale


def train(samples, params, epochs, lr=1e-3, log=print, val=(), lr_final=None):
    """Adam on the captured samples (one sample -- one System state of P packings -- per
    step, in a seeded random order each epoch; learning rate decayed geometrically to
    lr_final).  Returns (params, train losses, validation losses)."""
    Vmax = max(d["V"] for d in list(samples) + list(val))
    ctx = samples[0]["sys"].ctx
    net = Network(ctx, params, Vmax, train=True)
    rng = np.random.default_rng(0)
    losses, vlosses = [], []
    t0 = time.perf_counter()
    nsteps = max(1, epochs * len(samples))
    k = 0
    for ep in range(epochs):
        order = rng.permutation(len(samples))
        ep_loss = []
        for i in order:
            d = samples[i]
            a = lr if lr_final is None else lr * (lr_final / lr) ** (k / nsteps)
            ep_loss.append(net.train_step(d["V"], d["feat"].ptr, d["adj"].ptr, d["target"].ptr, a, want_loss=True))
            k += 1
        losses.append(float(np.mean(ep_loss)))
        if val:
            vlosses.append(float(np.mean([net.loss(d["V"], d["feat"].ptr, d["adj"].ptr, d["target"].ptr)
                                          for d in val])))
        log(f"[warmstart] epoch {ep}: train loss {losses[-1]:.5f}"
            + (f", validation {vlosses[-1]:.5f}" if val else "") + f" ({time.perf_counter() - t0:.1f} s)")
    return net.get_params(), losses, vlosses


def ab_compare(weights, packings, maxSteps=12000, applications=1, at_step=0, gamma=1.0, log=print, dtMax=0.1,
               dt0=1e-3):
    """Plain FIRE against FIRE with the learned displacement applied (one System each, same
    packings).  The learned run does `at_step` plain FIRE steps first (0 = a warm start of
    the initial packing), then applies the network `applications` times (each after a fresh
    evaluation, displacement scaled by gamma), then FIRE to convergence from a fresh FIRE
    state.  Both start FIRE at time step dt0.  Steps per packing count both FIRE phases;
    seconds include the network."""
    ell = packing_length_units(packings)
    s0, prm = fire_system(packings, dtMax=dtMax)
    conv0, steps0, t0 = fire_run(s0, prm, maxSteps, dt0=dt0)
    E0 = s0.get_energy()
    F0 = s0.max_abs_force()
    s1, prm1 = fire_system(packings, dtMax=dtMax)
    ws = WarmStart(weights)
    ws.outScale *= gamma
    pre, tpre = np.zeros(s1.P, int), 0.0
    if at_step > 0:
        c, n1, tpre = fire_run(s1, prm1, at_step, dt0=dt0)
        pre[:] = at_step
        pre[c >= 0] = c[c >= 0]
    tw = 0.0
    for _ in range(applications):
        s1.force_eval()
        tw += ws.apply(s1, ell)
    conv1, steps1, t1 = fire_run(s1, prm1, maxSteps, dt0=dt0)
    conv1 = np.where(conv1 >= 0, conv1 + pre, -1)
    E1 = s1.get_energy()
    F1 = s1.max_abs_force()
    res = dict(packings=len(packings), at_step=at_step, gamma=gamma, applications=applications, dt0=dt0,
               plain_steps=conv0.tolist(), warm_steps=conv1.tolist(),
               plain_seconds=t0, warm_seconds=tpre + t1 + tw, warm_apply_seconds=tw,
               plain_energy=E0.tolist(), warm_energy=E1.tolist(), plain_maxF=F0.tolist(), warm_maxF=F1.tolist())
    ok0, ok1 = conv0 >= 0, conv1 >= 0
    log(f"[warmstart] A/B on {len(packings)} packings (dt0 {dt0:g}, learned at step {at_step}, gamma {gamma}): "
        f"plain mean steps "
        f"{conv0[ok0].mean() if ok0.any() else -1:.0f} ({ok0.sum()} converged, {t0:.2f} s); learned mean steps "
        f"{conv1[ok1].mean() if ok1.any() else -1:.0f} ({ok1.sum()} converged, {tpre + t1 + tw:.2f} s incl. "
        f"{tw:.3f} s network)")
    return res
