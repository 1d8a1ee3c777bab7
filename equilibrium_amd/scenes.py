"""Seeded synthetic scenes for the benchmark workloads and the parity tests.

All inputs are generated on the host (never on the device) from a SplitMix64
stream, so the GPU run and the CPU oracle consume identical f32 arrays.  Each
scene is a dict: ``size`` (ring included), ``dt``, ``diffusion``,
``viscosity``, ``iters`` (Gauss-Seidel sweeps per lin_solve), ``frames``,
``rects`` [(x0, y0, x1, y1), half-open like fill_obstacle], per-frame sources
``src_d``/``src_u``/``src_v`` (row-major f32 of size*size, or None) and
optional initial velocity ``init_u``/``init_v``.

Sources are applied before every frame with add_density / add_velocity
semantics (fluid.rs:120-131), as the reference's frame loop applies add_noise
before step (renderer_helpers.rs:55-59); has_perlin_noise is off.
"""
import math

import numpy as np

MASK64 = (1 << 64) - 1
BASE_SEED = 0x5EED_F1D5


class SplitMix64:
    def __init__(self, seed):
        self.state = seed & MASK64

    def next_u64(self):
        self.state = (self.state + 0x9E3779B97F4A7C15) & MASK64
        z = self.state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK64
        return z ^ (z >> 31)

    def uniform(self, lo=0.0, hi=1.0):
        return lo + (hi - lo) * ((self.next_u64() >> 11) * (1.0 / (1 << 53)))

    def randint(self, lo, hi):
        """Integer in [lo, hi] inclusive."""
        return lo + self.next_u64() % (hi - lo + 1)

    def log_uniform(self, lo, hi):
        return math.exp(self.uniform(math.log(lo), math.log(hi)))

    def normal(self):
        u1 = max(self.uniform(), 1e-300)
        u2 = self.uniform()
        return math.sqrt(-2.0 * math.log(u1)) * math.cos(2.0 * math.pi * u2)


def derive_seed(*keys):
    r = SplitMix64(BASE_SEED)
    s = r.next_u64()
    for k in keys:
        r = SplitMix64(s ^ (int(k) * 0x9E3779B97F4A7C15 & MASK64))
        s = r.next_u64()
    return s


def gaussian(size, cx, cy, sigma, amp):
    c = np.arange(size, dtype=np.float64)
    r2 = ((c - cx) ** 2)[None, :] + ((c - cy) ** 2)[:, None]   # [y, x]
    g = amp * np.exp(-r2 / (2.0 * sigma * sigma))
    return g.astype(np.float32).ravel()


def _overlap(a, b, pad=1):
    return not (a[2] + pad <= b[0] or b[2] + pad <= a[0] or a[3] + pad <= b[1] or b[3] + pad <= a[1])


def place_rects(rng, N, count, lo_frac, hi_frac, avoid_points=(), tries=200):
    """Axis-aligned rectangles with sides U(lo_frac*N, hi_frac*N) at random
    interior positions, not overlapping each other nor the source points."""
    rects = []
    for _ in range(count):
        for _t in range(tries):
            w = max(1, int(round(rng.uniform(lo_frac * N, hi_frac * N))))
            h = max(1, int(round(rng.uniform(lo_frac * N, hi_frac * N))))
            if w >= N - 2 or h >= N - 2:
                continue
            x0 = rng.randint(2, N - w)
            y0 = rng.randint(2, N - h)
            r = (x0, y0, x0 + w, y0 + h)
            if any(_overlap(r, q) for q in rects):
                continue
            if any(r[0] - 2 <= px < r[2] + 2 and r[1] - 2 <= py < r[3] + 2 for px, py in avoid_points):
                continue
            rects.append(r)
            break
    return rects


def stam_scene(N, seed, *, dt=0.1, diffusion=0.0, viscosity=1e-4, iters=20, frames=100,
               n_rects=1, rect_frac=(0.1, 0.25), n_density=1, n_velocity=1, amp=100.0):
    """Config-0 style scene scaled to N: Gaussian density sources (sigma=N/16,
    amplitude `amp`) and constant-vector velocity patches (N/16 wide,
    |v|~U(1,5), angle~U(0,2pi)) at random interior points, plus rectangles."""
    rng = SplitMix64(seed)
    S = N + 2
    src_d = np.zeros(S * S, np.float32)
    src_u = np.zeros(S * S, np.float32)
    src_v = np.zeros(S * S, np.float32)
    points = []
    sigma = N / 16.0
    for _ in range(n_density):
        cx = rng.randint(1, N)
        cy = rng.randint(1, N)
        points.append((cx, cy))
        src_d += gaussian(S, cx, cy, sigma, amp)
    patch = max(2, N // 16)
    for _ in range(n_velocity):
        mag = rng.uniform(1.0, 5.0)
        ang = rng.uniform(0.0, 2.0 * math.pi)
        px = rng.randint(1, max(1, N - patch + 1))
        py = rng.randint(1, max(1, N - patch + 1))
        points.append((px + patch // 2, py + patch // 2))
        u = np.zeros((S, S), np.float32)
        v = np.zeros((S, S), np.float32)
        u[py:py + patch, px:px + patch] = np.float32(mag * math.cos(ang))
        v[py:py + patch, px:px + patch] = np.float32(mag * math.sin(ang))
        src_u += u.ravel()
        src_v += v.ravel()
    rects = place_rects(rng, N, n_rects, rect_frac[0], rect_frac[1], points)
    return dict(size=S, dt=dt, diffusion=diffusion, viscosity=viscosity, iters=iters,
                frames=frames, rects=rects, src_d=src_d, src_u=src_u, src_v=src_v,
                init_u=None, init_v=None)


def small_scene(seed=None):
    """N=128, dt 0.1, diffusion 0, viscosity 1e-4, 20 sweeps, 100 frames."""
    return stam_scene(128, derive_seed(0) if seed is None else seed)


def single_1024(seed=None):
    """N=1024, viscosity 1e-4, diffusion 1e-5, 20 sweeps, 200 frames,
    3 rectangles with sides U(N/20, N/8), 4 Gaussian density/velocity sources."""
    return stam_scene(1024, derive_seed(1) if seed is None else seed, diffusion=1e-5, frames=200,
                      n_rects=3, rect_frac=(1 / 20, 1 / 8), n_density=4, n_velocity=4)


def curl_noise(S, seed, smooth, peak):
    """Divergence-free velocity: curl of Gaussian-smoothed seeded noise.  The
    smoothing is periodic (a product with the Gaussian's transfer function
    exp(-2 pi^2 sigma^2 |k|^2) in the Fourier domain, on the next FFT-friendly
    size), about a second at S = 4098."""
    import scipy.fft

    g = np.random.Generator(np.random.PCG64(seed & MASK64))
    L = scipy.fft.next_fast_len(S, real=True)   # noise on an L x L torus, cropped to S x S
    k = np.fft.fftfreq(L)
    kr = np.fft.rfftfreq(L)
    h = np.exp(-2.0 * np.pi ** 2 * smooth ** 2 * (k[:, None] ** 2 + kr[None, :] ** 2))
    psi = scipy.fft.irfft2(scipy.fft.rfft2(g.standard_normal((L, L))) * h, s=(L, L))[:S, :S]
    u = np.zeros((S, S))
    v = np.zeros((S, S))
    u[1:-1, :] = (psi[2:, :] - psi[:-2, :]) * 0.5     # d psi / dy
    v[:, 1:-1] = -(psi[:, 2:] - psi[:, :-2]) * 0.5    # -d psi / dx
    mag = np.sqrt(u * u + v * v).max()
    scale = peak / mag if mag > 0 else 0.0
    return (u * scale).astype(np.float32).ravel(), (v * scale).astype(np.float32).ravel()


def redblack_scene(seed=None, N=4096):
    """N=4096, 40 sweeps, 100 frames, 16 rectangles, divergence-free curl-noise
    initial velocity with peak |u| = 2; dt 0.1, viscosity 1e-4, diffusion 1e-5
    (not fixed by the workload description); 4 Gaussian density sources."""
    seed = derive_seed(2) if seed is None else seed
    sc = stam_scene(N, seed, diffusion=1e-5, iters=40, frames=100, n_rects=16,
                    rect_frac=(1 / 40, 1 / 16), n_density=4, n_velocity=0)
    sc["src_u"] = None
    sc["src_v"] = None
    sc["init_u"], sc["init_v"] = curl_noise(N + 2, seed, smooth=max(2.0, N / 64.0), peak=2.0)
    return sc


def batched_mix(seed=None, n=512, frames=200):
    """Batched mix: N uniform in {64,128,256,512}, viscosity and diffusion
    log-uniform in [1e-6, 1e-3], 0-4 rectangles, 20 sweeps, dt 0.1."""
    base = derive_seed(3) if seed is None else seed
    rng = SplitMix64(base)
    out = []
    for i in range(n):
        N = (64, 128, 256, 512)[rng.randint(0, 3)]
        visc = rng.log_uniform(1e-6, 1e-3)
        diff = rng.log_uniform(1e-6, 1e-3)
        nr = rng.randint(0, 4)
        out.append(stam_scene(N, derive_seed(3, i, base), viscosity=visc, diffusion=diff,
                              frames=frames, n_rects=nr))
    return out


def serving_requests(rate, duration_s, seed=None, n_max=2048, n_min=64):
    """Serving workload: Poisson arrivals at `rate` scenes/s over `duration_s`;
    N = round(log-uniform[64, 2048]); frames U(50, 500); scenes like small_scene."""
    base = derive_seed(4, int(rate)) if seed is None else seed
    rng = SplitMix64(base)
    t = 0.0
    reqs = []
    i = 0
    while True:
        t += -math.log(max(rng.uniform(), 1e-300)) / rate
        if t > duration_s:
            break
        N = int(round(rng.log_uniform(n_min, n_max)))
        frames = rng.randint(50, 500)
        reqs.append((t, N, frames, derive_seed(4, i, base)))
        i += 1
    return reqs


def apply_to_scene(scene_handle, sc):
    """Marks obstacles, uploads initial velocity and registers frame sources
    on an _abi.Scene."""
    from . import _abi

    for r in sc["rects"]:
        scene_handle.fill_rect(*r)
    if sc.get("init_u") is not None:
        scene_handle.upload(_abi.VX, sc["init_u"])
        scene_handle.upload(_abi.VY, sc["init_v"])
    for f, key in ((_abi.DENSITY, "src_d"), (_abi.VX, "src_u"), (_abi.VY, "src_v")):
        if sc.get(key) is not None:
            scene_handle.set_frame_source(f, sc[key])


def scene_params(sc, gs_mode=0):
    from . import _abi

    p = _abi.default_params()
    p.size = sc["size"]
    p.delta_t = sc["dt"]
    p.frames = sc["frames"]
    p.diffusion = sc["diffusion"]
    p.viscosity = sc["viscosity"]
    p.gs_iters = sc["iters"]
    p.pressure_iters = sc["iters"]
    p.gs_mode = gs_mode
    p.init_times = 1
    return p
