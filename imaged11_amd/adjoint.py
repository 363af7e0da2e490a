"""Gradient of the joint grain + detector refinement objective.

ImageD11 fits detector geometry by Nelder-Mead on ``refinegrains.gof``
(``ImageD11/refinegrains.py:517-624``, ``ImageD11/simplex.py:51-95``), one
Python loop over grains per objective evaluation.  Here the objective

    L(UBIs, translations, geometry) =
        sum_g sum_{q in list g} || UBI_g . G_q(geometry, t_g) - hkl_q ||^2

with G_q = compute_gv(compute_xlylzl(sc, fc), omega) for grain g's
translation and assignments/hkl held fixed, is evaluated together with its
exact gradient in one pass over the list entries (``gs_joint_adjoint``: a
hand-written reverse sweep through compute_gv and compute_xlylzl per
(grain, peak) pair) and one launch that adds up each grain's pieces.  `JointLoss` exposes it as a ``torch.autograd.Function`` of the
UBIs, the translations and the 26 packed geometry numbers; `pack_geometry`
maps the physical parameters (distance, y/z centre, tilts, wavelength,
wedge, chi) onto that packing with differentiable torch operations, so
autograd carries the gradient to them.
"""
import ctypes
import math

import numpy as np
import torch

from . import _lib, batched, geometry
from ._lib import dptr, stream_ptr

GEO = 26
FREE = ("distance", "y_center", "z_center", "tilt_x", "tilt_y", "tilt_z", "wavelength", "wedge",
        "chi")


def pack_geometry(params, fixed, parity=False):
    """26-vector [p(4), r(9), dist(3), wavelength, mat(9)] from a dict of
    0-d tensors ``params`` (any of FREE) and plain numbers ``fixed`` (the
    rest of an ImageD11 parameter dict: pixel sizes, o11..o22, ...).

    Same conventions as transform.Ctransform (``ImageD11/transform.py:
    685-693``) and the wedge/chi matrix of compute_gv
    (``src/cdiffraction.c:119-142``); differentiable in ``params``.
    parity=True gives mat the values of the reference's glibc sincos
    (``batched.wedge_chi_matrix``) while its gradient stays torch's."""
    q = geometry.full_parameters(fixed)
    dev = next(iter(params.values())).device if params else torch.device("cpu")

    def P(name):
        if name in params:
            return params[name]
        return torch.tensor(float(q[name]), dtype=torch.float64, device=dev)

    one = torch.ones((), dtype=torch.float64, device=dev)
    zero = torch.zeros((), dtype=torch.float64, device=dev)
    tx, ty, tz = P("tilt_x"), P("tilt_y"), P("tilt_z")
    rz = torch.stack([torch.stack([torch.cos(tz), -torch.sin(tz), zero]),
                      torch.stack([torch.sin(tz), torch.cos(tz), zero]),
                      torch.stack([zero, zero, one])])
    ry = torch.stack([torch.stack([torch.cos(ty), zero, torch.sin(ty)]),
                      torch.stack([zero, one, zero]),
                      torch.stack([-torch.sin(ty), zero, torch.cos(ty)])])
    rx = torch.stack([torch.stack([one, zero, zero]),
                      torch.stack([zero, torch.cos(tx), -torch.sin(tx)]),
                      torch.stack([zero, torch.sin(tx), torch.cos(tx)])])
    flip = torch.tensor(geometry.flip_matrix(q["o11"], q["o12"], q["o21"], q["o22"]),
                        dtype=torch.float64, device=dev)
    r = (rx @ ry @ rz @ flip).reshape(9)
    p = torch.stack([P("z_center"), P("y_center"), P("z_size"), P("y_size")])
    dist = torch.stack([P("distance"), zero, zero])
    w = P("wedge") * (np.pi / 180.0)
    c = P("chi") * (np.pi / 180.0)
    W = torch.stack([torch.stack([torch.cos(w), zero, -torch.sin(w)]),
                     torch.stack([zero, one, zero]),
                     torch.stack([torch.sin(w), zero, torch.cos(w)])])
    C = torch.stack([torch.stack([one, zero, zero]),
                     torch.stack([zero, torch.cos(c), -torch.sin(c)]),
                     torch.stack([zero, torch.sin(c), torch.cos(c)])])
    mat = (W @ C).reshape(9)
    if parity:
        exact = torch.as_tensor(batched.wedge_chi_matrix(float(P("wedge").detach()),
                                                         float(P("chi").detach())),
                                dtype=torch.float64, device=dev).reshape(9)
        mat = mat + (exact - mat).detach()
    return torch.cat([p, r, dist, P("wavelength").reshape(1), mat])


class JointProblem:
    """Device-resident data of a joint refinement: the raw peak table
    (sc, fc and the per-peak sin/cos of omega), per-grain CSR peak lists and
    the fixed integer hkl of every list entry.

    When every list index is a valid peak number, (sc, fc, sin, cos) are
    copied once per list entry in list order, so each evaluation streams
    them instead of gathering four cache lines per entry; with an
    out-of-range index the lists stay indexed (such entries are skipped).
    Either way the per-peak data are those of the construction.

    parity=True takes sin/cos of omega from the host's glibc, the values
    the reference's compute_gv uses (``batched.sincos_omega``); otherwise
    the device's correctly rounded sincos.

    Each evaluation runs over pieces of at most ``piece_len`` consecutive
    list entries (``gs_joint_schedule``, longest first), so a long list is
    shared by several waves instead of setting the pass's duration."""

    def __init__(self, sc, fc, omega, omegasign, idx, off, parity=False, piece_len=512):
        batched._need(sc, torch.float64, "sc", (None,))
        n = sc.shape[0]
        batched._need(fc, torch.float64, "fc", (n,))
        batched._need(omega, torch.float64, "omega", (n,))
        batched._need(idx, torch.int32, "idx", (None,))
        batched._need(off, torch.int64, "off", (None,))
        self.sc, self.fc = sc, fc
        self.so, self.co = batched.sincos_omega(omega, omegasign, parity=parity)
        self.parity = bool(parity)
        self.idx, self.off = idx, off
        self.rows = None
        m = idx.numel()
        if m == 0 or (int(idx.min()) >= 0 and int(idx.max()) < n):
            il = idx.long()
            self.rows = tuple(a[il].contiguous() for a in (sc, fc, self.so, self.co))
        self.K = off.numel() - 1
        lengths = off[1:] - off[:-1]
        self.owner = torch.repeat_interleave(
            torch.arange(self.K, dtype=torch.int32, device=sc.device), lengths)
        self.hkl = torch.zeros((idx.numel(), 3), dtype=torch.int32, device=sc.device)
        dev = sc.device
        self.piece_len = int(piece_len)
        self.piece_off, self.piece_grain, self.slots = schedule(off, idx.numel(), self.piece_len)
        self.n_pieces = self.slots.numel()
        lib = _lib.load()
        self.ws = torch.empty(max(1, int(lib.gs_joint_workspace_doubles(self.n_pieces))),
                              dtype=torch.float64, device=dev)
        self.loss_g = torch.empty(self.K, dtype=torch.float64, device=dev)

    def assign_hkl(self, geo, ubis, trans):
        """Fix hkl_q = round(UBI_g . G_q) at the given state."""
        lib = _lib.load()
        lib.gs_clear_error()
        # bound to names so that no argument's memory is recycled by the
        # allocator before the launch (a temporary's block is free again as
        # soon as its last reference goes); geo may come from the host
        geo = geo.detach().to(self.sc.device, torch.float64).contiguous()
        ubis = ubis.detach().contiguous()
        trans = trans.detach().contiguous()
        (a, b, c, d), npk, idx = self._peaks()
        _lib.check(lib.gs_joint_hkl(dptr(a), dptr(b), dptr(c), dptr(d), npk, dptr(geo),
                                    dptr(ubis), dptr(trans), self.K, idx, dptr(self.owner),
                                    self.idx.numel(), dptr(self.hkl), stream_ptr()), "joint hkl")

    def _peaks(self):
        """(sc, fc, sin, cos), their length and the index pointer the
        library takes: the per-entry copies without an index, or the peak
        table with it."""
        if self.rows is not None:
            return self.rows, self.idx.numel(), None
        return (self.sc, self.fc, self.so, self.co), self.sc.shape[0], dptr(self.idx)

    def _static_args(self):
        """The launch arguments that do not change between evaluations
        (formed again if the per-entry copies are switched off)."""
        cached = getattr(self, "_args", None)
        if cached is not None and cached[0] is self.rows:
            return cached[1], cached[2]
        (a, b, c, d), npk, idx = self._peaks()
        head = (dptr(a), dptr(b), dptr(c), dptr(d), npk)
        mid = (self.K, idx, self.idx.numel(), dptr(self.off), dptr(self.hkl),
               dptr(self.piece_off), dptr(self.piece_grain), dptr(self.slots), self.n_pieces,
               self.piece_len)
        self._args = (self.rows, head, mid)
        return head, mid

    def evaluate(self, geo, ubis, trans, gspace=False, want_sums=False, lanes_per_piece=16,
                 geo_part=None):
        """One fused forward + reverse pass.  Returns (loss, dL/dUBI (K,3,3),
        dL/dt (K,3), dL/dgeo (26,), sums) with sums = (K,15) Gauss-Newton
        sums or None.  In g-space mode the loss is 1/2 sum ||UB.h - G||^2
        and the UBI gradient slot holds dE/dUB.  geo_part: an optional
        (K,26) tensor that receives each grain's share of dL/dgeo.  The
        outputs are views of one new allocation per call."""
        lib = _lib.load()
        dev = self.sc.device
        geo, ubis, trans = (_dev64(x, dev) for x in (geo, ubis, trans))
        K = self.K
        n = 12 * K + GEO + 1 + (15 * K if want_sums else 0)
        buf = torch.empty(n, dtype=torch.float64, device=dev)
        gu = buf[:9 * K].view(K, 3, 3)
        gt = buf[9 * K:12 * K].view(K, 3)
        out = buf[12 * K:12 * K + GEO + 1]
        sums = buf[12 * K + GEO + 1:].view(K, 15) if want_sums else None
        if geo_part is not None:
            batched._need(geo_part, torch.float64, "geo_part", (K, GEO))
        head, mid = self._static_args()
        _lib.check(lib.gs_joint_adjoint(
            *head, dptr(geo), dptr(ubis), dptr(trans), *mid, int(bool(gspace)),
            dptr(self.loss_g), dptr(gu), dptr(gt),
            dptr(geo_part) if geo_part is not None else None,
            dptr(sums) if want_sums else None, dptr(out), int(lanes_per_piece),
            dptr(self.ws), stream_ptr()), "joint adjoint")
        return out[GEO], gu, gt, out[:GEO], sums


def _dev64(x, dev):
    """x as a contiguous float64 tensor on dev without a gradient (the
    tensor itself when it already is one)."""
    x = x.detach()
    if x.device != dev or x.dtype != torch.float64:
        x = x.to(dev, torch.float64)
    return x if x.is_contiguous() else x.contiguous()


def schedule(off, n_idx, piece_len):
    """The pieces of a pass (``gs_joint_schedule``, host C): device tensors
    piece_off (K+1, int64), piece_grain and slots (int32)."""
    lib = _lib.load()
    lib.gs_clear_error()
    dev = off.device
    off_h = off.detach().to("cpu", torch.int64).contiguous()
    K = off_h.numel() - 1
    piece_off = torch.empty(K + 1, dtype=torch.int64)
    n = ctypes.c_longlong(0)
    _lib.check(lib.gs_joint_schedule(off_h.data_ptr(), K, int(n_idx), int(piece_len),
                                     piece_off.data_ptr(), None, None, 0, ctypes.byref(n)),
               "joint schedule")
    grain = torch.empty(max(1, n.value), dtype=torch.int32)
    slots = torch.empty(max(1, n.value), dtype=torch.int32)
    _lib.check(lib.gs_joint_schedule(off_h.data_ptr(), K, int(n_idx), int(piece_len),
                                     piece_off.data_ptr(), grain.data_ptr(), slots.data_ptr(),
                                     n.value, ctypes.byref(n)), "joint schedule")
    return (piece_off.to(dev), grain[:n.value].to(dev), slots[:n.value].to(dev))


class JointLoss(torch.autograd.Function):
    """L(UBIs, translations, packed geometry) with its hand-written adjoint."""

    @staticmethod
    def forward(ctx, ubis, trans, geo, problem):
        loss, gu, gt, gg, _ = problem.evaluate(geo, ubis, trans)
        ctx.geo_device = geo.device
        ctx.save_for_backward(gu, gt, gg)
        return loss

    @staticmethod
    def backward(ctx, grad):
        gu, gt, gg = ctx.saved_tensors
        return grad * gu, grad * gt, (grad * gg).to(ctx.geo_device), None


def joint_loss(problem, ubis, trans, params, fixed):
    """L as a differentiable function of UBIs, translations and the
    physical geometry parameters in ``params`` (dict of 0-d tensors)."""
    geo = pack_geometry(params, fixed, parity=problem.parity)
    return JointLoss.apply(ubis, trans, geo, problem)


def gauss_newton_ubis(sums, ubis=None, out=None):
    """Per-grain least-squares UBI for fixed g-vectors and hkl:
    UBI = (sum h G^T)(sum G G^T)^-1 from the (K,15) sums of an evaluation
    (``gs_gauss_newton_ubis``, one device thread per grain, cofactor
    inverse).  A grain whose sum G G^T is singular keeps its UBI from
    ``ubis`` (as score_and_refine does; NaN when ubis is None).  out: a
    (K,3,3) float64 device tensor to write (may be ``ubis``), else a new
    one."""
    lib = _lib.load()
    K = sums.shape[0]
    batched._need(sums, torch.float64, "sums", (K, 15))
    if ubis is not None:
        batched._need(ubis, torch.float64, "ubis", (K, 3, 3))
    if out is None:
        out = torch.empty((K, 3, 3), dtype=torch.float64, device=sums.device)
    batched._need(out, torch.float64, "out", (K, 3, 3))
    _lib.check(lib.gs_gauss_newton_ubis(dptr(sums), dptr(ubis) if ubis is not None else None,
                                        K, dptr(out), stream_ptr()), "gauss_newton_ubis")
    return out


def _rot(axis, a):
    """Rotation by a about axis 0/1/2 with the sign conventions of
    geometry.detector_rotation_matrix, and its derivative d/da."""
    c, s = math.cos(a), math.sin(a)
    if axis == 0:
        m = [[1.0, 0.0, 0.0], [0.0, c, -s], [0.0, s, c]]
        d = [[0.0, 0.0, 0.0], [0.0, -s, -c], [0.0, c, -s]]
    elif axis == 1:
        m = [[c, 0.0, s], [0.0, 1.0, 0.0], [-s, 0.0, c]]
        d = [[-s, 0.0, c], [0.0, 0.0, 0.0], [-c, 0.0, -s]]
    else:
        m = [[c, -s, 0.0], [s, c, 0.0], [0.0, 0.0, 1.0]]
        d = [[-s, -c, 0.0], [c, -s, 0.0], [0.0, 0.0, 0.0]]
    return np.array(m), np.array(d)


# slots of the packed geometry that depend on one parameter alone, linearly
_LINEAR_SLOT = {"z_center": 0, "y_center": 1, "z_size": 2, "y_size": 3, "distance": 13,
                "wavelength": 16}


def pack_geometry_host(values, fixed, parity=False):
    """``pack_geometry`` on the host without autograd: the 26 packed
    numbers for the plain-number parameters ``values`` (names of FREE, or
    z_size/y_size; the rest from ``fixed``) and their derivatives in closed
    form, as (geo (26,), jac (26, len(values)) in the order of ``values``).
    Tilts enter r = Rx.Ry.Rz.flip, wedge and chi (degrees) mat = W.C
    (``ImageD11/transform.py:685-693``, ``src/cdiffraction.c:119-142``);
    parity=True takes mat's values from the host's glibc sincos as
    pack_geometry does."""
    q = geometry.full_parameters(fixed)
    names = list(values)
    for k in names:
        q[k] = float(values[k])
    flip = geometry.flip_matrix(q["o11"], q["o12"], q["o21"], q["o22"])
    rx, dx = _rot(0, q["tilt_x"])
    ry, dy = _rot(1, q["tilt_y"])
    rz, dz = _rot(2, q["tilt_z"])
    deg = np.pi / 180.0
    w, c = q["wedge"] * deg, q["chi"] * deg
    cw, sw, cc, sc = math.cos(w), math.sin(w), math.cos(c), math.sin(c)
    W = np.array([[cw, 0.0, -sw], [0.0, 1.0, 0.0], [sw, 0.0, cw]])
    dW = np.array([[-sw, 0.0, -cw], [0.0, 0.0, 0.0], [cw, 0.0, -sw]]) * deg
    C = np.array([[1.0, 0.0, 0.0], [0.0, cc, -sc], [0.0, sc, cc]])
    dC = np.array([[0.0, 0.0, 0.0], [0.0, -sc, -cc], [0.0, cc, -sc]]) * deg
    geo = np.empty(GEO)
    geo[0:4] = (q["z_center"], q["y_center"], q["z_size"], q["y_size"])
    geo[4:13] = (rx @ ry @ rz @ flip).ravel()
    geo[13:16] = (q["distance"], 0.0, 0.0)
    geo[16] = q["wavelength"]
    geo[17:26] = (batched.wedge_chi_matrix(q["wedge"], q["chi"]) if parity else W @ C).ravel()
    jac = np.zeros((GEO, len(names)))
    for j, k in enumerate(names):
        if k in _LINEAR_SLOT:
            jac[_LINEAR_SLOT[k], j] = 1.0
        elif k == "tilt_x":
            jac[4:13, j] = (dx @ ry @ rz @ flip).ravel()
        elif k == "tilt_y":
            jac[4:13, j] = (rx @ dy @ rz @ flip).ravel()
        elif k == "tilt_z":
            jac[4:13, j] = (rx @ ry @ dz @ flip).ravel()
        elif k == "wedge":
            jac[17:26, j] = (dW @ C).ravel()
        elif k == "chi":
            jac[17:26, j] = (W @ dC).ravel()
        else:
            raise ValueError("pack_geometry_host: %r is not a geometry parameter" % (k,))
    return geo, jac


class Adam:
    """Adam on a float64 vector (Kingma & Ba 2015, with bias correction),
    in the arithmetic of ``torch.optim.Adam``'s default single-tensor
    update: m <- m + (1 - b1)(g - m), v <- b2 v + (1 - b2) g^2,
    x <- x - lr/(1 - b1^t) . m / (sqrt(v)/sqrt(1 - b2^t) + eps)."""

    def __init__(self, n, lr, betas=(0.9, 0.999), eps=1e-8):
        self.lr, self.b1, self.b2, self.eps = float(lr), betas[0], betas[1], eps
        self.m = np.zeros(n)
        self.v = np.zeros(n)
        self.t = 0

    def step(self, x, g):
        """x after one step with gradient g (a new array)."""
        self.t += 1
        self.m += (1.0 - self.b1) * (g - self.m)
        self.v = self.v * self.b2 + (1.0 - self.b2) * (g * g)
        step_size = self.lr / (1.0 - self.b1 ** self.t)
        denom = np.sqrt(self.v) / math.sqrt(1.0 - self.b2 ** self.t) + self.eps
        return x - step_size * (self.m / denom)


def refine_joint(problem, ubis, trans, params, fixed, steps=50, lr=1e-3, scales=None):
    """Joint refinement: each step is one adjoint pass; the UBIs take the
    Gauss-Newton (exact least-squares) update for the current geometry
    (``gs_gauss_newton_ubis``, in place on the device) and the geometry
    parameters an Adam step on dL/dparams (per-parameter step scales in
    ``scales``).  hkl and the peak lists stay fixed.  The handful of
    geometry numbers live on the host: the 26 packed values and their
    Jacobian come from ``pack_geometry_host``, go to the device once per
    step, and dL/dgeo with the loss come back in one copy, the step's only
    synchronisation.  Returns (ubis, params, loss history)."""
    dev = problem.sc.device
    ubis = _dev64(ubis, dev).clone()
    trans = _dev64(trans, dev)
    names = sorted(params)
    scale = np.array([float((scales or {}).get(k, 1.0)) for k in names])
    # optimise dimensionless offsets x_k with value = start + scale_k * x_k
    start = np.array([float(params[k]) for k in names])
    x = np.zeros(len(names))
    opt = Adam(len(names), lr)
    pinned = dev.type == "cuda"
    geo_h = torch.empty(GEO, dtype=torch.float64, pin_memory=pinned)
    back_h = torch.empty(GEO + 1, dtype=torch.float64, pin_memory=pinned)
    geo_d = torch.empty(GEO, dtype=torch.float64, device=dev)
    stream = torch.cuda.current_stream(dev) if pinned else None
    history = []
    for _ in range(steps):
        cur = start + scale * x
        geo, jac = pack_geometry_host(dict(zip(names, cur)), fixed, parity=problem.parity)
        geo_h.numpy()[:] = geo
        geo_d.copy_(geo_h, non_blocking=True)
        _, _, _, gg, sums = problem.evaluate(geo_d, ubis, trans, want_sums=True)
        gauss_newton_ubis(sums, ubis, out=ubis)
        # gg (26) and the loss are adjacent in the evaluation's buffer
        back_h.copy_(gg.as_strided((GEO + 1,), (1,)), non_blocking=True)
        if stream is not None:
            stream.synchronize()
        back = back_h.numpy()
        history.append(float(back[GEO]))
        x = opt.step(x, (jac.T @ back[:GEO]) * scale)
    final = {k: torch.tensor(float(v), dtype=torch.float64, device=params[k].device)
             for k, v in zip(names, start + scale * x)}
    return ubis, final, history
