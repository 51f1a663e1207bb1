"""Case descriptions, `.cfg` text and the seeded synthetic workloads.

`Case` mirrors the reference's CaseConfig (include/gsc/config.hpp:41-67) plus
the extension fields of gsc_case_desc. `to_cfg()` writes a `.cfg` file the
reference reader (src/config.cpp:243-347) accepts, always with explicit map_*
keys (a file without them silently keeps the default [-2,2]^2 map,
config.cpp:343-345) and doubles printed with %.17g so they round-trip exactly.

The generators build the benchmark workloads:
  unicycle_30x30x24x10   30x30x24x10 grid x 25 inputs, two seeded obstacles
  unicycle_91x105x42x10  91x105x42x10 grid x 100 inputs, hyperbolic target band
  batch instances        log-uniform 50k-1M cells, 16/36/64/100 inputs
  unicycle_128x128x64x16 16.8M cells x 100 inputs
  serving requests       log-uniform 20k-4M cells
Every shape is checked against the reference's ceil((ub-lb)/eta) rule
(grid.hpp:121, 157), which can add a cell when the quotient rounds up.
"""
from __future__ import annotations

import ctypes
import math
from dataclasses import dataclass, field
from typing import List, Optional, Sequence, Tuple

import numpy as np

from . import capi

PI_LITERAL = 3.141592653589793
Rect = Tuple[Tuple[float, float, float, float], Tuple[float, float, float, float]]


def _r(lb, ub) -> Rect:
    return (tuple(float(v) for v in lb), tuple(float(v) for v in ub))


@dataclass
class Case:
    name: str = "unicycle_default"
    state_lb: Tuple[float, ...] = (-2.0, -2.0, -PI_LITERAL, 0.0)
    state_ub: Tuple[float, ...] = (2.0, 2.0, PI_LITERAL, 2.0)
    state_eta: Tuple[float, ...] = (0.5, 0.5, 0.7853981633974483, 0.5)
    wrap_dim: int = 2
    input_lb: Tuple[float, ...] = (-0.5, -1.0)
    input_ub: Tuple[float, ...] = (0.5, 1.0)
    input_eta: Tuple[float, ...] = (0.5, 1.0)
    candidate: Rect = ((-0.5, -0.5, -PI_LITERAL, 0.0), (0.5, 0.5, PI_LITERAL, 1.5))
    map: Rect = ((-2.0, -2.0, -PI_LITERAL, 0.0), (2.0, 2.0, PI_LITERAL, 2.0))
    obstacles: List[Rect] = field(default_factory=list)
    constraint_type: int = capi.CONSTRAINT_NONE
    hyperbolic_a: float = 4.0
    hyperbolic_b: float = 4.0
    hyperbolic_c: float = 16.0
    elliptic_a: float = 2.0
    elliptic_b: float = 2.0
    disturbance_half_width: Tuple[float, ...] = (0.05, 0.05, 0.02, 0.02)
    sample_time: float = 0.2
    integration_substeps: int = 32
    max_iterations: int = 32
    verbose: bool = False
    extra_targets: List[Rect] = field(default_factory=list)
    target_band: Optional[Tuple[float, float]] = None

    # ---- conversion to the C descriptor ----------------------------------
    def to_desc(self):
        """Return (CaseDesc, keepalive) — keep `keepalive` referenced during calls."""
        d = capi.CaseDesc()
        name = self.name.encode()
        d.name = name
        d.state_lb[:] = self.state_lb
        d.state_ub[:] = self.state_ub
        d.state_eta[:] = self.state_eta
        d.wrap_dim = int(self.wrap_dim)
        d.input_lb[:] = self.input_lb
        d.input_ub[:] = self.input_ub
        d.input_eta[:] = self.input_eta
        d.candidate.lb[:] = self.candidate[0]
        d.candidate.ub[:] = self.candidate[1]
        d.map.lb[:] = self.map[0]
        d.map.ub[:] = self.map[1]
        obs = (capi.Rect4 * max(1, len(self.obstacles)))()
        for k, (lb, ub) in enumerate(self.obstacles):
            obs[k].lb[:] = lb
            obs[k].ub[:] = ub
        d.obstacles = ctypes.cast(obs, ctypes.POINTER(capi.Rect4)) if self.obstacles else None
        d.obstacle_count = len(self.obstacles)
        d.constraint_type = int(self.constraint_type)
        d.hyperbolic_a, d.hyperbolic_b, d.hyperbolic_c = self.hyperbolic_a, self.hyperbolic_b, self.hyperbolic_c
        d.elliptic_a, d.elliptic_b = self.elliptic_a, self.elliptic_b
        d.disturbance_half_width[:] = self.disturbance_half_width
        d.sample_time = self.sample_time
        d.integration_substeps = int(self.integration_substeps)
        d.max_iterations = int(self.max_iterations)
        d.verbose = 1 if self.verbose else 0
        tg = (capi.Rect4 * max(1, len(self.extra_targets)))()
        for k, (lb, ub) in enumerate(self.extra_targets):
            tg[k].lb[:] = lb
            tg[k].ub[:] = ub
        d.extra_targets = ctypes.cast(tg, ctypes.POINTER(capi.Rect4)) if self.extra_targets else None
        d.extra_target_count = len(self.extra_targets)
        if self.target_band is not None:
            d.target_band_enabled = 1
            d.target_band_c, d.target_band_w = self.target_band
        return d, (name, obs, tg)

    @staticmethod
    def from_desc(d: capi.CaseDesc) -> "Case":
        rects = lambda ptr, n: [_r(ptr[k].lb, ptr[k].ub) for k in range(n)]
        return Case(
            name=(d.name or b"").decode(),
            state_lb=tuple(d.state_lb),
            state_ub=tuple(d.state_ub),
            state_eta=tuple(d.state_eta),
            wrap_dim=d.wrap_dim,
            input_lb=tuple(d.input_lb),
            input_ub=tuple(d.input_ub),
            input_eta=tuple(d.input_eta),
            candidate=_r(d.candidate.lb, d.candidate.ub),
            map=_r(d.map.lb, d.map.ub),
            obstacles=rects(d.obstacles, d.obstacle_count) if d.obstacle_count else [],
            constraint_type=d.constraint_type,
            hyperbolic_a=d.hyperbolic_a,
            hyperbolic_b=d.hyperbolic_b,
            hyperbolic_c=d.hyperbolic_c,
            elliptic_a=d.elliptic_a,
            elliptic_b=d.elliptic_b,
            disturbance_half_width=tuple(d.disturbance_half_width),
            sample_time=d.sample_time,
            integration_substeps=d.integration_substeps,
            max_iterations=d.max_iterations,
            verbose=bool(d.verbose),
            extra_targets=rects(d.extra_targets, d.extra_target_count) if d.extra_target_count else [],
            target_band=(d.target_band_c, d.target_band_w) if d.target_band_enabled else None,
        )

    # ---- .cfg text ---------------------------------------------------------
    def to_cfg(self) -> str:
        f = lambda xs: ", ".join("%.17g" % v for v in xs)
        ct = {capi.CONSTRAINT_NONE: "none", capi.CONSTRAINT_HYPERBOLIC: "hyperbolic",
              capi.CONSTRAINT_ELLIPTIC: "elliptic"}[self.constraint_type]
        lines = [
            f"name = {self.name}",
            f"state_lb = {f(self.state_lb)}",
            f"state_ub = {f(self.state_ub)}",
            f"state_eta = {f(self.state_eta)}",
            f"wrap_dim = {self.wrap_dim}",
            f"input_lb = {f(self.input_lb)}",
            f"input_ub = {f(self.input_ub)}",
            f"input_eta = {f(self.input_eta)}",
            f"candidate_lb = {f(self.candidate[0])}",
            f"candidate_ub = {f(self.candidate[1])}",
            f"map_lb = {f(self.map[0])}",
            f"map_ub = {f(self.map[1])}",
            f"constraint_type = {ct}",
            f"hyperbolic_a = {self.hyperbolic_a:.17g}",
            f"hyperbolic_b = {self.hyperbolic_b:.17g}",
            f"hyperbolic_c = {self.hyperbolic_c:.17g}",
            f"elliptic_a = {self.elliptic_a:.17g}",
            f"elliptic_b = {self.elliptic_b:.17g}",
            f"disturbance_half_width = {f(self.disturbance_half_width)}",
            f"sample_time = {self.sample_time:.17g}",
            f"integration_substeps = {self.integration_substeps}",
            f"max_iterations = {self.max_iterations}",
            f"verbose = {'true' if self.verbose else 'false'}",
        ]
        for lb, ub in self.obstacles:
            lines.append(f"obstacle = {f(tuple(lb) + tuple(ub))}")
        for lb, ub in self.extra_targets:
            lines.append(f"target = {f(tuple(lb) + tuple(ub))}")
        if self.target_band is not None:
            lines.append(f"target_band = {f(self.target_band)}")
        return "\n".join(lines) + "\n"

    def reference_cfg(self) -> str:
        """The `.cfg` without extension keys (what the reference reader accepts)."""
        return "\n".join(l for l in self.to_cfg().splitlines() if not l.startswith(("target =", "target_band ="))) + "\n"

    def shape(self) -> Tuple[Tuple[int, ...], Tuple[int, int]]:
        """Grid shapes by the reference rule ceil((ub - lb) / eta) (grid.hpp:121, 157)."""
        s = tuple(int(math.ceil((self.state_ub[d] - self.state_lb[d]) / self.state_eta[d])) for d in range(4))
        i = tuple(int(math.ceil((self.input_ub[d] - self.input_lb[d]) / self.input_eta[d])) for d in range(2))
        return s, i

    def state_count(self) -> int:
        s, _ = self.shape()
        return s[0] * s[1] * s[2] * s[3]

    def input_count(self) -> int:
        _, i = self.shape()
        return i[0] * i[1]


def _config_from_handle(lib, handle) -> Case:
    try:
        return Case.from_desc(lib.gsc_config_desc(handle).contents)
    finally:
        lib.gsc_config_free(handle)


def load_cfg(path: str) -> Case:
    """Read a `.cfg` file with the library's reader (same rules as config.cpp:243-347)."""
    lib = capi.load()
    h = ctypes.c_void_p()
    err = ctypes.create_string_buffer(512)
    rc = lib.gsc_config_load(path.encode(), ctypes.byref(h), err, len(err))
    if rc != capi.GSC_OK:
        raise ValueError(err.value.decode())
    return _config_from_handle(lib, h)


def parse_cfg(text: str) -> Case:
    lib = capi.load()
    h = ctypes.c_void_p()
    err = ctypes.create_string_buffer(512)
    rc = lib.gsc_config_parse(text.encode(), ctypes.byref(h), err, len(err))
    if rc != capi.GSC_OK:
        raise ValueError(err.value.decode())
    return _config_from_handle(lib, h)


BUILTIN = {"default": 0, "wraparound": 1, "smoke": 2, "hyperbolic": 3}


def builtin(which: str) -> Case:
    """The reference's canned cases (config.cpp:129-241)."""
    lib = capi.load()
    h = ctypes.c_void_p()
    rc = lib.gsc_config_builtin(BUILTIN[which], ctypes.byref(h))
    if rc != capi.GSC_OK:
        raise ValueError(which)
    return _config_from_handle(lib, h)


# ---------------------------------------------------------------------------
# Seeded generators
# ---------------------------------------------------------------------------
def fit_eta(width: float, n: int) -> float:
    """Largest-error-free eta with ceil(width / eta) == n under IEEE double."""
    eta = width / n
    for _ in range(64):
        k = math.ceil(width / eta)
        if k == n:
            return eta
        eta = math.nextafter(eta, math.inf) if k > n else math.nextafter(eta, 0.0)
    raise ValueError(f"cannot fit {n} cells in width {width}")


def _domain(shape, xy=(0.0, 10.0), v=(0.0, 2.0)):
    lb = (xy[0], xy[0], -PI_LITERAL, v[0])
    ub = (xy[1], xy[1], PI_LITERAL, v[1])
    eta = tuple(fit_eta(ub[d] - lb[d], shape[d]) for d in range(4))
    return lb, ub, eta


def _inputs(n_side: int):
    lb = (-1.0, -1.0)
    ub = (1.0, 1.0)
    eta = (fit_eta(2.0, n_side), fit_eta(2.0, n_side))
    return lb, ub, eta


def _xy_overlap(a: Rect, b: Rect) -> bool:
    return not (a[1][0] < b[0][0] or b[1][0] < a[0][0] or a[1][1] < b[0][1] or b[1][1] < a[0][1])


def _draw_obstacles(rng, count, avoid: Sequence[Rect], lb, ub, side=(0.5, 2.0), band=None):
    out = []
    while len(out) < count:
        cx, cy = rng.uniform(lb[0], ub[0]), rng.uniform(lb[1], ub[1])
        sx, sy = rng.uniform(*side), rng.uniform(*side)
        r = _r((cx - sx / 2, cy - sy / 2, lb[2], lb[3]), (cx + sx / 2, cy + sy / 2, ub[2], ub[3]))
        if any(_xy_overlap(r, a) for a in avoid):
            continue
        if band is not None:
            c, w = band
            x0, y0 = max(r[0][0], 0.0), max(r[0][1], 0.0)
            x1, y1 = max(r[1][0], 0.0), max(r[1][1], 0.0)
            if x0 * y0 <= c + w and x1 * y1 >= c - w:
                continue
        out.append(r)
    return out


def hyperbolic_demo() -> Case:
    """The reference's published case, examples/hyperbolic_case.cfg: a 70x70x63x13
    grid (ceil(1.2/0.1) rounds up to 13) x 10x10 inputs = 401,310,000 pairs under
    the hyperbolic state constraint (README.md:44-102)."""
    lb = (-3.5, -3.5, -PI_LITERAL, -0.1)
    ub = (3.5, 3.5, PI_LITERAL, 1.1)
    return Case(
        name="hyperbolic_demo", state_lb=lb, state_ub=ub, state_eta=(0.1, 0.1, 0.1, 0.1), wrap_dim=2,
        input_lb=(-1.0, -1.0), input_ub=(1.0, 1.0), input_eta=(0.2, 0.2),
        candidate=_r((-0.8, -0.8, -PI_LITERAL, 0.0), (0.8, 0.8, PI_LITERAL, 1.0)), map=_r(lb, ub),
        constraint_type=capi.CONSTRAINT_HYPERBOLIC, hyperbolic_a=4.0, hyperbolic_b=4.0, hyperbolic_c=16.0,
        disturbance_half_width=(0.05, 0.05, 0.05, 0.05), sample_time=0.5, integration_substeps=32,
        max_iterations=100, verbose=True,
    )


def unicycle_30x30x24x10(sample_time: float = 0.1, seed: int = 0) -> Case:
    """30x30x24x10 grid over x,y in [0,10], θ in [-π,π), v in [0,2]; 5x5 inputs
    (a, ω) in [-1,1]^2; target x,y in [4,6]; two obstacles from `seed`."""
    lb, ub, eta = _domain((30, 30, 24, 10))
    ilb, iub, ieta = _inputs(5)
    target = _r((4.0, 4.0, -PI_LITERAL, 0.0), (6.0, 6.0, PI_LITERAL, 2.0))
    rng = np.random.default_rng(seed)
    obstacles = _draw_obstacles(rng, 2, [target], lb, ub)
    return Case(
        name=f"unicycle_30x30x24x10_T{sample_time:g}_s{seed}",
        state_lb=lb, state_ub=ub, state_eta=eta, input_lb=ilb, input_ub=iub, input_eta=ieta,
        candidate=target, map=_r(lb, ub), obstacles=obstacles, sample_time=sample_time,
        integration_substeps=32, max_iterations=200,
    )


def unicycle_91x105x42x10(sample_time: float = 1.0, seed: int = 1) -> Case:
    """91x105x42x10 grid (4,013,100 states) x 10x10 inputs = 401,310,000 pairs.
    Target: cells of the safe box with |x*y - c| <= w, (c, w) from `seed`;
    three obstacles from `seed` kept off the band."""
    lb, ub, eta = _domain((91, 105, 42, 10))
    ilb, iub, ieta = _inputs(10)
    rng = np.random.default_rng(seed)
    c = float(rng.uniform(15.0, 35.0))
    w = float(rng.uniform(1.5, 3.0))
    safe = _r((0.4, 0.4, -PI_LITERAL, 0.0), (9.6, 9.6, PI_LITERAL, 2.0))
    obstacles = _draw_obstacles(rng, 3, [], safe[0], safe[1], band=(c, w))
    return Case(
        name=f"unicycle_91x105x42x10_s{seed}",
        state_lb=lb, state_ub=ub, state_eta=eta, input_lb=ilb, input_ub=iub, input_eta=ieta,
        candidate=safe, map=safe, obstacles=obstacles, sample_time=sample_time,
        integration_substeps=32, max_iterations=200, target_band=(c, w),
    )


def unicycle_128x128x64x16(sample_time: float = 0.5, seed: int = 4) -> Case:
    """128x128x64x16 grid (16,777,216 states) x 100 inputs; target x,y in [4,6]
    and two obstacles from `seed`, as in the 30x30x24x10 case."""
    lb, ub, eta = _domain((128, 128, 64, 16))
    ilb, iub, ieta = _inputs(10)
    target = _r((4.0, 4.0, -PI_LITERAL, 0.0), (6.0, 6.0, PI_LITERAL, 2.0))
    rng = np.random.default_rng(seed)
    obstacles = _draw_obstacles(rng, 2, [target], lb, ub)
    return Case(
        name=f"unicycle_128x128x64x16_s{seed}",
        state_lb=lb, state_ub=ub, state_eta=eta, input_lb=ilb, input_ub=iub, input_eta=ieta,
        candidate=target, map=_r(lb, ub), obstacles=obstacles, sample_time=sample_time,
        integration_substeps=32, max_iterations=200,
    )


def random_instance(rng: np.random.Generator, cells_lo: float, cells_hi: float, name: str,
                    sample_time: float = 0.5) -> Case:
    """One batch/serving instance: cells log-uniform in [cells_lo, cells_hi] with
    per-axis extents ~9:10:4:1, inputs from {16, 36, 64, 100}, 1-3 target boxes
    and 0-5 obstacle boxes."""
    cells = math.exp(rng.uniform(math.log(cells_lo), math.log(cells_hi)))
    k = (cells / 360.0) ** 0.25
    shape = (max(2, round(9 * k)), max(2, round(10 * k)), max(2, round(4 * k)), max(2, round(k)))
    lb, ub, eta = _domain(shape)
    side = int(rng.choice([4, 6, 8, 10]))
    ilb, iub, ieta = _inputs(side)
    n_t = int(rng.integers(1, 4))
    targets = []
    for _ in range(n_t):
        cx, cy = rng.uniform(1.5, 8.5), rng.uniform(1.5, 8.5)
        sx, sy = rng.uniform(1.0, 2.5), rng.uniform(1.0, 2.5)
        targets.append(_r((cx - sx / 2, cy - sy / 2, -PI_LITERAL, 0.0), (cx + sx / 2, cy + sy / 2, PI_LITERAL, 2.0)))
    n_o = int(rng.integers(0, 6))
    obstacles = _draw_obstacles(rng, n_o, targets, lb, ub)
    return Case(
        name=name, state_lb=lb, state_ub=ub, state_eta=eta, input_lb=ilb, input_ub=iub, input_eta=ieta,
        candidate=targets[0], map=_r(lb, ub), obstacles=obstacles, sample_time=sample_time,
        integration_substeps=32, max_iterations=200, extra_targets=targets[1:],
    )


def batch_instances(count: int = 32, seed: int = 3) -> List[Case]:
    rng = np.random.default_rng(seed)
    return [random_instance(rng, 5e4, 1e6, f"batch_{seed}_{k}") for k in range(count)]


def serving_requests(count: int, seed: int = 5) -> List[Case]:
    rng = np.random.default_rng(seed)
    return [random_instance(rng, 2e4, 4e6, f"request_{seed}_{k}") for k in range(count)]
