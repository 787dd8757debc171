"""Float64 reference of the batched frame forward for gradient checks of the
adjoint ray-cast (test infrastructure).

The sample set of every pixel -- which blocks it visits, front to back, and
which samples of each segment are composited before the segment or the pixel
terminates -- is taken from the oracle (orc_ray_samples, the fp32 march of
renderArgb.hpp:39-51 with the frame's acc_cutoff), together with each
sample's trilinear corners and weights.  With those decisions fixed, the
image is a smooth function of the colour table and the brick values, which
this module evaluates in float64 with torch so that it can be differentiated
(autograd) and finite-differenced:

    v_s = trilinear(brick, corners_s, weights_s)
    t_s = (1 - f) tf[i0] + f tf[i1],  x = (v_s - fmin) / (fmax - fmin) * N - 0.5
          (clamped to [-1, N], i0 = clamp(floor x), i1 = clamp(floor x + 1))
    a_s = 1 - (1 - t_s.w * opacity_factor) ^ i_layer_count
    P = sum_s T_s a_s t_s.rgb,  T_s = prod_{s' < s} (1 - a_s'),  A = 1 - prod_s (1 - a_s)
    image = (P / A, A)

(DenseFieldRenderer_cu.cu:62-215, renderArgb.hpp:30-55, cuda_blend.hpp:24-33,
DenseFieldRenderer_cu.cu:573-589 for the colour-table fetch.)
"""
import numpy as np

from revise_amd import api


class SampleTable:
    """Kept samples of the pixels (vx, vy) of a crop, in compositing order."""

    def __init__(self, orc, blocks, bricks_np, M, W, H, fov, fmin, fmax, thr, quality, cutoff,
                 acc_cutoff, tf_np, crop, max_samples=2048):
        """max_samples: the longest segment (samples) taken whole."""
        x0, y0, cw, ch = crop
        keys = np.array([(b["level"],) + tuple(b["index"]) for b in blocks], dtype=np.uint32)
        f2b = api.mixed_back_to_front(keys, api.eye_from_view(M))[::-1]
        po = {}
        for i in f2b:
            b = blocks[int(i)]
            live, p = orc.block_params(M, W, H, fov, b["level"], b["index"], b["depth"], fmin, fmax, thr,
                                       quality, cutoff)
            if live:
                po[int(i)] = p
        tf_np = np.ascontiguousarray(tf_np, dtype=np.float32)
        pix, blk, corners, weights, order = [], [], [], [], []
        self.npix = cw * ch
        for py in range(ch):
            for px in range(cw):
                vx, vy = x0 + px, y0 + py
                acc = np.float32(0.0)
                k = 0
                for i in f2b:
                    i = int(i)
                    p = po.get(i)
                    if p is None or not (p.view_x <= vx < p.view_x + p.view_w and
                                         p.view_y <= vy < p.view_y + p.view_h):
                        continue
                    c, w, a = orc.ray_samples(p, bricks_np[i], tf_np, vx, vy, max_samples=max_samples)
                    if c.shape[0]:
                        pix.append(np.full(c.shape[0], py * cw + px, dtype=np.int64))
                        blk.append(np.full(c.shape[0], i, dtype=np.int64))
                        corners.append(c.astype(np.int64))
                        weights.append(w.astype(np.float64))
                        order.append(np.arange(k, k + c.shape[0], dtype=np.int64))
                        k += c.shape[0]
                    # the frame's pixel alpha (blendBehind of the segment), in fp32
                    da = np.float32(a * (np.float32(1.0) - acc))
                    acc = np.float32(acc + da)
                    if acc > np.float32(acc_cutoff):
                        break
        cat = lambda xs, dt: np.concatenate(xs) if xs else np.zeros(0, dtype=dt)
        self.pix = cat(pix, np.int64)
        self.blk = cat(blk, np.int64)
        self.corners = cat(corners, np.int64).reshape(-1, 8)
        self.weights = cat(weights, np.float64).reshape(-1, 3)
        self.order = cat(order, np.int64)
        self.L = int(self.order.max()) + 1 if self.order.size else 1
        self.touched = np.unique(self.blk)
        der = {i: orc.block_derived(po[i]) for i in po}
        self.of = {i: float(der[i][3]) for i in po}
        self.ilc = {i: float(der[i][2]) for i in po}
        self.fmin, self.fmax = float(fmin), float(fmax)
        self.crop = crop

    def forward(self, torch, tf64, bricks64, device):
        """Crop image (ch, cw, 4), rows indexed by vy - y0 (bottom-up), float64.
        bricks64: (nblocks, V) float64 tensor (only touched rows are read)."""
        t = lambda a, dt=torch.int64: torch.as_tensor(a, dtype=dt, device=device)
        blk, cor, w = t(self.blk), t(self.corners), t(self.weights, torch.float64)
        vals = bricks64[blk[:, None], cor]                       # (S, 8)
        fx, fy, fz = w[:, 0], w[:, 1], w[:, 2]
        lerp = lambda a, b, f: (1 - f) * a + f * b
        c00 = lerp(vals[:, 0], vals[:, 1], fx)
        c10 = lerp(vals[:, 2], vals[:, 3], fx)
        c01 = lerp(vals[:, 4], vals[:, 5], fx)
        c11 = lerp(vals[:, 6], vals[:, 7], fx)
        v = lerp(lerp(c00, c10, fy), lerp(c01, c11, fy), fz)
        N = tf64.shape[0]
        x = ((v - self.fmin) / (self.fmax - self.fmin)) * N - 0.5
        x = torch.clamp(x, -1.0, float(N))
        fl = torch.floor(x)
        f = (x - fl)[:, None]
        i = fl.long()
        i0 = torch.clamp(i, 0, N - 1)
        i1 = torch.clamp(i + 1, 0, N - 1)
        tcol = (1 - f) * tf64[i0] + f * tf64[i1]                   # (S, 4)
        of = t(np.array([self.of[int(b)] for b in self.blk]), torch.float64)
        ilc = t(np.array([self.ilc[int(b)] for b in self.blk]), torch.float64)
        alpha = 1 - torch.pow(1 - tcol[:, 3] * of, ilc)
        npix = self.npix
        pix, order = t(self.pix), t(self.order)
        Apad = torch.zeros((npix, self.L), dtype=torch.float64, device=device)
        Apad = Apad.index_put((pix, order), alpha)
        Cpad = torch.zeros((npix, self.L, 3), dtype=torch.float64, device=device)
        Cpad = Cpad.index_put((pix, order), tcol[:, :3])
        keep = torch.cumprod(1 - Apad, dim=1)
        T = torch.cat([torch.ones((npix, 1), dtype=torch.float64, device=device), keep[:, :-1]], dim=1)
        P = ((T * Apad)[..., None] * Cpad).sum(1)
        A = 1 - keep[:, -1]
        safeA = torch.where(A > 0, A, torch.ones_like(A))
        rgb = torch.where((A > 0)[:, None], P / safeA[:, None], torch.zeros_like(P))
        x0, y0, cw, ch = self.crop
        return torch.cat([rgb, A[:, None]], dim=1).reshape(ch, cw, 4)
