"""Learned geometry warm start for calibration.

A small network predicts each view's geometry error from the measured
projections and the reprojection of the known volume at the nominal geometry;
its prediction is the starting point of the gradient calibration
(``calibration.calibrate``), and the plain start (the nominal geometry) is the
fallback whenever the prediction does not lower the reprojection residual.

Network (~2M parameters, all GEMMs on BF16 MFMA with FP32 accumulation, see
include/cbct_nn.h):
  input   per view a [128,128,2] image: the reprojection at the current geometry and
          the difference (measured - reprojection), both as 4x4 block means of the
          full-resolution projections, scaled to O(1)
  encoder five 3x3 stride-2 convolutions with ReLU, 2->32->64->128->256->256
          channels (128^2 -> 4^2), then the spatial mean: 256 features per view
  pooling per view: (its features, the mean of the features over the views)
  head    512 -> 768 -> 768 -> 8 (ReLU between), one correction per geometry
          parameter in units of the perturbation scale (``ScanFamily.scales``)
Training pairs come from the projector only: random ellipsoid phantoms, a
nominal circular scan, random per-view perturbations of random overall size;
the target is the perturbation.  The loss is the reprojection error of the
predicted correction, linearised: the mean squared detector displacement of the
volume's corners and centre (``cbct_nn_jacobian_mse``), so the network spends
no capacity on directions that do not move the projections.  Training runs on
the GPU with the network's own backward kernels and Adam
(``WarmStartNet.train_step``).

At use the prediction is applied a few times (``warm_start``): reproject at the
current estimate, predict the remaining correction, keep it only while the
reprojection residual falls; gradient calibration continues from there.
"""

import math

import numpy as np
import torch

from . import _native, calibration, nn_ops, phantom

GEOMETRY_NAMES = calibration.GEOMETRY_NAMES


class NetConfig:
    """Channel plan of the encoder convolutions and of the head (fc[0] = 2 * conv[-1])."""

    def __init__(self, conv=(2, 32, 64, 128, 256, 256), fc=(512, 768, 768, 8), image=128):
        self.conv = tuple(int(c) for c in conv)
        self.fc = tuple(int(c) for c in fc)
        self.image = int(image)
        if self.fc[0] != 2 * self.conv[-1]:
            raise ValueError("fc[0] must be twice the last conv width")
        if any(c % 8 for c in self.conv[1:]) or any(c % 8 for c in self.fc):
            raise ValueError("layer widths (except the input channels) must be multiples of 8")

    @staticmethod
    def kpad(c_in):
        return -(-(9 * c_in + 1) // 8) * 8

    def param_shapes(self):
        shapes = []
        for i in range(len(self.conv) - 1):
            shapes.append((f"conv{i + 1}", (self.conv[i + 1], self.kpad(self.conv[i]))))
        for i in range(len(self.fc) - 1):
            shapes.append((f"fc{i + 1}_w", (self.fc[i + 1], self.fc[i])))
            shapes.append((f"fc{i + 1}_b", (self.fc[i + 1],)))
        return shapes

    def feature_size(self):
        s = self.image
        for _ in range(len(self.conv) - 1):
            s = (s - 1) // 2 + 1
        return s


class WarmStartNet:
    """Parameters live in one flat FP32 buffer (segments 8-element aligned) with a bf16 copy
    used by the GEMMs, a gradient buffer and the Adam moments."""

    def __init__(self, cfg=None, seed=0, device="cuda"):
        self.cfg = cfg or NetConfig()
        self.device = torch.device(device)
        self.offsets = {}
        off = 0
        for name, shape in self.cfg.param_shapes():
            n = int(np.prod(shape))
            self.offsets[name] = (off, shape)
            off += -(-n // 8) * 8
        self.size = off
        self.theta = torch.tensor(self._init(seed), dtype=torch.float32, device=self.device)
        self.theta_b = nn_ops.to_bf16(self.theta)
        self.grad = torch.zeros_like(self.theta)
        self.m = torch.zeros_like(self.theta)
        self.v = torch.zeros_like(self.theta)
        self.t = 0
        self.cache = None

    def _init(self, seed):
        """He-normal weights on the real taps, zero biases and padding; the last layer 10x smaller."""
        rng = np.random.default_rng(seed)
        flat = np.zeros(self.size, dtype=np.float64)
        nfc = len(self.cfg.fc) - 1
        for name, (off, shape) in self.offsets.items():
            w = np.zeros(shape)
            if name.startswith("conv"):
                i = int(name[4:]) - 1
                cin = self.cfg.conv[i]
                w[:, :9 * cin] = rng.standard_normal((shape[0], 9 * cin)) * math.sqrt(2.0 / (9 * cin))
            elif name.endswith("_w"):
                w = rng.standard_normal(shape) * math.sqrt(2.0 / shape[1])
                if name == f"fc{nfc}_w":
                    w *= 0.1
            flat[off:off + w.size] = w.reshape(-1)
        return flat

    @property
    def num_parameters(self):
        return sum(int(np.prod(shape)) for _, shape in self.offsets.values())

    def param(self, name, buf=None):
        off, shape = self.offsets[name]
        buf = self.theta if buf is None else buf
        return buf[off:off + int(np.prod(shape))].view(shape)

    def params_numpy(self):
        return {name: self.param(name).double().cpu().numpy() for name in self.offsets}

    def grads_numpy(self):
        return {name: self.param(name, self.grad).double().cpu().numpy() for name in self.offsets}

    def load_params(self, params):
        flat = np.zeros(self.size, dtype=np.float32)
        for name, (off, shape) in self.offsets.items():
            flat[off:off + int(np.prod(shape))] = np.asarray(params[name], np.float32).reshape(-1)
        self.theta.copy_(torch.from_numpy(flat).to(self.device))
        nn_ops.to_bf16(self.theta, self.theta_b)

    def save(self, path):
        np.savez(path, **{k: v.astype(np.float32) for k, v in self.params_numpy().items()})

    def load(self, path):
        with np.load(path, allow_pickle=False) as f:
            self.load_params({k: f[k] for k in f.files})

    # ------------------------------------------------------------------ forward / backward
    def forward(self, x, P, A, keep=False):
        """x [P*A, image, image, conv[0]] bf16 -> predictions [P*A, fc[-1]] fp32."""
        cfg = self.cfg
        if x.dtype != nn_ops.BF16 or x.shape != (P * A, cfg.image, cfg.image, cfg.conv[0]):
            raise ValueError(f"input must be bf16 of shape {(P * A, cfg.image, cfg.image, cfg.conv[0])}")
        cache = {"P": P, "A": A, "conv": [], "fc": []}
        h = x.contiguous()
        nconv = len(cfg.conv) - 1
        for i in range(nconv):
            Wb = self.param(f"conv{i + 1}", self.theta_b)
            co, kp = Wb.shape
            N, H, W, _ = h.shape
            Ho, Wo = (H - 1) // 2 + 1, (W - 1) // 2 + 1
            col = nn_ops.im2col(h, 2, kp)
            M = N * Ho * Wo
            y = torch.empty((N, Ho, Wo, co), dtype=torch.float32, device=x.device)
            last = i == nconv - 1
            yb = None if last else torch.empty((N, Ho, Wo, co), dtype=nn_ops.BF16, device=x.device)
            nn_ops.gemm(col, Wb, M, co, kp, 0, 1, kp, kp, C=y, Cb=yb, relu=True)
            if keep:
                cache["conv"].append((tuple(h.shape), col, y))
            h = yb
        NI, S, C = P * A, y.shape[1] * y.shape[2], y.shape[3]
        feat = nn_ops.spatial_mean(y, NI, S, C)
        a = nn_ops.view_concat(feat, P, A, C)
        cache["pool"] = (S, C, tuple(y.shape))
        nfc = len(cfg.fc) - 1
        for i in range(nfc):
            Wb = self.param(f"fc{i + 1}_w", self.theta_b)
            b = self.param(f"fc{i + 1}_b")
            out_f, in_f = Wb.shape
            last = i == nfc - 1
            z = torch.empty((NI, out_f), dtype=torch.float32, device=x.device)
            zb = None if last else torch.empty((NI, out_f), dtype=nn_ops.BF16, device=x.device)
            nn_ops.gemm(a, Wb, NI, out_f, in_f, 0, 1, in_f, in_f, C=z, Cb=zb, bias=b, relu=not last)
            if keep:
                cache["fc"].append((a, z))
            a = zb if not last else z
        self.cache = cache if keep else None
        return a

    def backward(self, dout):
        """Parameter gradients of sum(dout * forward(x)) into ``self.grad`` (overwritten)."""
        cache = self.cache
        if cache is None:
            raise RuntimeError("backward needs a forward(..., keep=True) first")
        cfg = self.cfg
        P, A = cache["P"], cache["A"]
        NI = P * A
        nfc = len(cfg.fc) - 1
        d = dout.contiguous()
        for i in reversed(range(nfc)):
            a_in, z = cache["fc"][i]
            Wb = self.param(f"fc{i + 1}_w", self.theta_b)
            out_f, in_f = Wb.shape
            if i == nfc - 1:
                dz, dzb = d, nn_ops.to_bf16(d)
            else:
                dz, dzb = nn_ops.relu_backward(d, z)
            nn_ops.gemm(dzb, a_in, out_f, in_f, NI, 1, 0, out_f, in_f, C=self.param(f"fc{i + 1}_w", self.grad))
            nn_ops.colsum(dz, NI, out_f, self.param(f"fc{i + 1}_b", self.grad))
            d = torch.empty((NI, in_f), dtype=torch.float32, device=dout.device)
            nn_ops.gemm(dzb, Wb, NI, in_f, out_f, 0, 0, out_f, in_f, C=d)
        S, C, yshape = cache["pool"]
        dfeat = nn_ops.view_concat_backward(d, P, A, C)
        dy = nn_ops.spatial_mean_backward(dfeat, NI, S, C).view(yshape)
        nconv = len(cfg.conv) - 1
        for i in reversed(range(nconv)):
            in_shape, col, y = cache["conv"][i]
            Wb = self.param(f"conv{i + 1}", self.theta_b)
            co, kp = Wb.shape
            M = col.shape[0]
            dz, dzb = nn_ops.relu_backward(dy, y)
            tiles = -(-co // 64) * -(-kp // 64)
            splits = max(1, min(-(-M // 2048), 1024 // tiles))
            nn_ops.gemm(dzb, col, co, kp, M, 1, 0, co, kp, C=self.param(f"conv{i + 1}", self.grad), splits=splits)
            if i > 0:
                dcol = torch.empty((M, kp), dtype=torch.float32, device=dout.device)
                nn_ops.gemm(dzb, Wb, M, kp, co, 0, 0, co, kp, C=dcol)
                dy, _ = nn_ops.col2im(dcol, in_shape, 2, kp)
        self.cache = None

    def train_step(self, x, target, P, A, lr, beta1=0.9, beta2=0.999, eps=1e-8, weight_decay=0.0, jacobian=None):
        """One Adam step on the mean squared error, or with ``jacobian`` ([Q,K] or [P*A,Q,K])
        on the linearised reprojection loss; returns the loss tensor (on the GPU)."""
        pred = self.forward(x, P, A, keep=True)
        if jacobian is None:
            loss, dpred = nn_ops.mse(pred, target)
        else:
            loss, dpred = nn_ops.jacobian_mse(pred, target, jacobian)
        self.backward(dpred)
        self.t += 1
        nn_ops.adam(self.theta, self.theta_b, self.grad, self.m, self.v, lr, beta1, beta2, eps,
                    1.0 - beta1 ** self.t, 1.0 - beta2 ** self.t, weight_decay)
        return loss


# ---------------------------------------------------------------------------- scans and data
class ScanFamily:
    """The scans the network is trained for: 256^3 ellipsoid phantoms (voxel 0.5 mm), A=180
    views of a full circle, 512x512 detector of 0.8 mm pixels, SOD 600 mm, SID 1100 mm, step
    0.25 mm; network images are 4x4 block means (128x128).  Perturbations as in the
    256^3 x 720 workload: angle N(0, 0.2 deg), centres N(0, 1.5 px), source N(0, 0.5 mm),
    SID and SOD x (1 + N(0, 0.01))."""

    def __init__(self, D=256, A=180, U=512, V=512, pixel=0.8, voxel=0.5, step=0.25, sod=600.0, sid=1100.0,
                 factor=4, diff_gain=4.0, angle_deg=0.2, center_px=1.5, src_mm=0.5, distance_rel=0.01):
        self.D, self.A, self.U, self.V = D, A, U, V
        self.pixel, self.voxel, self.step, self.sod, self.sid = pixel, voxel, step, sod, sid
        self.factor = factor
        self.diff_gain = diff_gain
        self.perturb = dict(angle_deg=angle_deg, center_px=center_px, src_mm=src_mm, distance_rel=distance_rel)
        # the perturbation scale of each geometry parameter (network output units)
        self.scales = torch.tensor([math.radians(angle_deg), distance_rel * sid, distance_rel * sod, center_px,
                                    center_px, src_mm, src_mm, src_mm], dtype=torch.float64)
        # input scale: line integrals over the volume width are O(D * voxel)
        self.input_scale = 1.0 / (D * voxel)

    def nominal(self, device):
        g = phantom.circular_geometry(self.A, self.sid, self.sod, self.U, self.V, self.pixel)
        return torch.stack([g[k] for k in GEOMETRY_NAMES]).to(device)

    def spec(self, seed):
        return {"D": self.D, "A": self.A, "U": self.U, "V": self.V, "pixel": self.pixel, "voxel": self.voxel,
                "sod": self.sod, "sid": self.sid, "step": self.step, "seed": int(seed)}

    def phantom(self, seed, device):
        params = phantom.ellipsoid_parameters(int(seed), 12, half_width=self.D * self.voxel / 2)
        return phantom.rasterize_ellipsoids_on(params, (self.D,) * 3, self.voxel, device)

    def perturbed(self, nominal, seed, size=1.0):
        """nominal + ``size`` times a random perturbation of the family's scales."""
        g = {k: nominal[i].cpu() for i, k in enumerate(GEOMETRY_NAMES)}
        p = phantom.perturb_geometry(g, int(seed), **{k: v * size for k, v in self.perturb.items()})
        return torch.stack([p[k] for k in GEOMETRY_NAMES]).to(nominal.device)

    def jacobian(self, geometry):
        """[A, 18, 8] fp32: detector displacement (u, v in pixels) of the volume's eight corners
        and centre per unit correction of each parameter (in units of ``scales``), by central
        differences of the backprojector's point-to-detector map in FP64."""
        g = geometry.detach().double().cpu()
        pts = calibration.fixed_points((self.D,) * 3, self.voxel)
        cols = []
        for k in range(8):
            h = 1e-3 * float(self.scales[k])
            gp, gm = g.clone(), g.clone()
            gp[k] += h
            gm[k] -= h
            up, vp = calibration.detector_positions(gp, pts, self.pixel)
            um, vm = calibration.detector_positions(gm, pts, self.pixel)
            du = (up - um) / (2 * h) * float(self.scales[k])
            dv = (vp - vm) / (2 * h) * float(self.scales[k])
            cols.append(torch.cat([du, dv], dim=1))             # [A, 18]
        return torch.stack(cols, dim=2).float().contiguous().to(geometry.device)

    def project(self, volume, geometry):
        return _native.fp_forward(volume, [g.contiguous() for g in geometry], self.pixel, self.U, self.V,
                                  self.step, self.voxel, 0.0, 0.0, 0.0)

    def network_input(self, measured, reprojection):
        return nn_ops.prepare_input(measured, reprojection, self.factor, self.input_scale, self.diff_gain)


class TrainingData:
    """Projector-generated training pairs.  A pool of phantoms with their nominal-geometry
    projections is kept on the GPU; each pair draws a phantom and a fresh perturbation and
    projects it (one forward projection per pair)."""

    def __init__(self, family, pool=16, seed=1000, device="cuda", size=(0.05, 1.3)):
        self.family = family
        self.device = torch.device(device)
        self.size = size
        self.nominal = family.nominal(self.device)
        self.J = family.jacobian(self.nominal)
        self.volumes, self.nominal_proj = [], []
        for i in range(pool):
            vol = family.phantom(seed + i, self.device)
            self.volumes.append(vol)
            self.nominal_proj.append(family.project(vol, self.nominal))
        self.rng = np.random.default_rng(seed)
        self.counter = seed * 7919

    def pair(self, index=None):
        """(network input [A,128,128,2] bf16, target [A,8] fp32 in units of the scales)."""
        f = self.family
        i = int(self.rng.integers(len(self.volumes))) if index is None else index
        self.counter += 1
        lo, hi = self.size
        size = math.exp(self.rng.uniform(math.log(lo), math.log(hi)))
        true = f.perturbed(self.nominal, self.counter, size)
        measured = f.project(self.volumes[i], true)
        x = f.network_input(measured, self.nominal_proj[i])
        target = ((true.double() - self.nominal.double()).T / f.scales.to(self.device)).float().contiguous()
        return x, target

    def batch(self, P):
        xs, ts = zip(*[self.pair() for _ in range(P)])
        return torch.cat(xs, 0), torch.cat(ts, 0)

    def jacobian_rows(self, P):
        return self.J.repeat(P, 1, 1).contiguous()


def train(net, data, steps, P=2, lr=1e-3, lr_final=1e-5, log_every=0, log=print):
    """Adam on the linearised reprojection loss with a cosine learning-rate decay from lr to
    lr_final; returns the loss history (mean squared detector displacement, px^2)."""
    A = data.family.A
    J = data.jacobian_rows(P)
    hist = []
    for s in range(steps):
        x, t = data.batch(P)
        rate = lr_final + 0.5 * (lr - lr_final) * (1 + math.cos(math.pi * s / max(1, steps - 1)))
        loss = net.train_step(x, t, P, A, rate, jacobian=J)
        hist.append(loss)
        if log_every and (s + 1) % log_every == 0:
            log(f"train step {s + 1}/{steps}: loss {loss.item():.4f}")
    return [h.item() for h in hist]


# ---------------------------------------------------------------------------- use
def predict_correction(net, family, problem, reprojection):
    """Geometry correction [8, A] (parameter units) from the measured projections and the
    reprojection at the problem's current geometry."""
    x = family.network_input(problem.measured, reprojection)
    A = problem.measured.shape[0]
    pred = net.forward(x, 1, A)
    return pred.double().T * family.scales.to(pred.device)[:, None]


def warm_start(problem, net, family, iterations=3):
    """Move ``problem.geometry`` to the network's estimate, applying the prediction up to
    ``iterations`` times (reproject, predict the remaining correction) and keeping each
    update only if the reprojection residual falls; the starting geometry is the fallback.
    Returns the residuals (mean squared, per accepted iterate, the start first)."""
    g = problem.geometry.clone()
    proj = family.project(problem.volume, g)
    residuals = [nn_ops.sqdiff_mean(proj, problem.measured).item()]
    for _ in range(iterations):
        problem.geometry.copy_(g)
        cand = (g.double() + predict_correction(net, family, problem, proj)).float()
        cproj = family.project(problem.volume, cand)
        r = nn_ops.sqdiff_mean(cproj, problem.measured).item()
        if not r < residuals[-1]:
            break
        g, proj = cand, cproj
        residuals.append(r)
    problem.geometry.copy_(g)
    return residuals
