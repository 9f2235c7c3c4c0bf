"""Synthetic cloth / soft-body scenes of the VBD workloads.

Each function returns ``(model, state, info)`` where ``state`` carries the
seeded initial positions and ``info`` the step parameters.  Shared choices:

* up axis Y, gravity -9.81, frame dt 1/60 s split into 10 substeps;
* cloth grids are built as ``add_cloth_grid`` builds them, rotated +90 deg
  about x so the grid lies in the XZ plane, with lumped mass from an areal
  density (``add_cloth_mesh``: density * area / 3 per triangle corner);
* contact material is averaged between ``model.soft_contact_*`` and each
  shape (arithmetic mean of ke / kd, geometric mean of mu), so the shape mu is
  chosen to give the stated effective friction: ground plane mu 0.5 with
  ``soft_contact_mu`` 0.5, sphere mu 0.18 for an effective 0.3.
"""

from __future__ import annotations

import math

import numpy as np
import torch

from .geometry import ParticleFlags
from .sim.builder import ModelBuilder, cloth_grid_mesh, quat_from_axis_angle

FRAME_DT = 1.0 / 60.0
SUBSTEPS = 10


def _finish(builder: ModelBuilder, device, jitter_sigma: float, seed: int, soft_ke: float, soft_kd: float,
            soft_mu: float):
    model = builder.finalize(device)
    model.soft_contact_ke = soft_ke
    model.soft_contact_kd = soft_kd
    model.soft_contact_mu = soft_mu
    state = model.state()
    if jitter_sigma > 0.0:
        rng = np.random.default_rng(seed)
        jitter = rng.normal(0.0, jitter_sigma, size=(model.particle_count, 3)).astype(np.float32)
        state.particle_q += torch.as_tensor(jitter, device=model.device)
    return model, state


def cloth_hanging(device="cuda:0", verts_per_side: int = 32, size: float = 1.0, height: float = 1.0,
                  seed: int = 0, iterations: int = 5):
    """32x32 cloth, two top corners pinned, ground plane (the 1,024-vertex CPU-runnable case).

    Areal density 0.1 kg/m^2, tri_ke = tri_ka = 1e4, tri_kd = 1e-5, edge_ke = 1,
    contact ke 1e4 / mu 0.5, positions jittered by N(0, 1e-4) from ``seed``.
    """
    n = verts_per_side - 1
    b = ModelBuilder(up_axis="Y", gravity=-9.81)
    b.add_ground_plane(cfg=ModelBuilder.ShapeConfig(ke=1.0e4, kd=0.0, mu=0.5))
    verts, tris = cloth_grid_mesh(n, n, size / n, size / n)
    rot = quat_from_axis_angle((1.0, 0.0, 0.0), math.pi / 2)
    b.add_cloth_mesh(pos=(-0.5 * size, height, -0.5 * size), rot=rot, scale=1.0, vel=(0.0, 0.0, 0.0),
                     vertices=verts, indices=tris, density=0.1, tri_ke=1.0e4, tri_ka=1.0e4, tri_kd=1.0e-5,
                     edge_ke=1.0, edge_kd=0.0)
    top = [n * (n + 1), n * (n + 1) + n]
    b.particle_mass[top] = 0.0
    b.particle_flags[top] = int(ParticleFlags.ACTIVE) & ~int(ParticleFlags.ACTIVE)
    b.color(include_bending=True)
    model, state = _finish(b, device, 1.0e-4, seed, 1.0e4, 0.0, 0.5)
    info = {"workload": f"cloth_hanging_{verts_per_side}x{verts_per_side}", "dt": FRAME_DT / SUBSTEPS,
            "substeps": SUBSTEPS, "iterations": iterations}
    return model, state, info


def cloth_drop_on_sphere(device="cuda:0", verts_per_side: int = 256, size: float = 2.0, height: float = 1.5,
                         seed: int = 1, iterations: int = 10, springs: bool = False, spring_ke: float = 100.0,
                         spring_kd: float = 1.0e-3):
    """256x256 cloth over 2 m falling onto a static sphere (r 0.4 m) and the ground (65,536 vertices).

    Same material as :func:`cloth_hanging` but tri_ke per triangle from
    log-uniform[5e3, 2e4] (``seed``); sphere contact ke 1e4, effective mu 0.3.
    With ``springs`` the cloth also gets a spring on every mesh edge and across
    every hinge (``add_cloth_mesh(add_springs=True)``, ke ``spring_ke``, kd ``spring_kd``).
    """
    n = verts_per_side - 1
    b = ModelBuilder(up_axis="Y", gravity=-9.81)
    b.add_ground_plane(cfg=ModelBuilder.ShapeConfig(ke=1.0e4, kd=0.0, mu=0.5))
    b.add_shape_sphere(-1, radius=0.4, cfg=ModelBuilder.ShapeConfig(ke=1.0e4, kd=0.0, mu=0.18))
    verts, tris = cloth_grid_mesh(n, n, size / n, size / n)
    rot = quat_from_axis_angle((1.0, 0.0, 0.0), math.pi / 2)
    b.add_cloth_mesh(pos=(-0.5 * size, height, -0.5 * size), rot=rot, scale=1.0, vel=(0.0, 0.0, 0.0),
                     vertices=verts, indices=tris, density=0.1, tri_ke=1.0e4, tri_ka=1.0e4, tri_kd=1.0e-5,
                     edge_ke=1.0, edge_kd=0.0, add_springs=springs, spring_ke=spring_ke, spring_kd=spring_kd)
    rng = np.random.default_rng(seed)
    ke = np.exp(rng.uniform(math.log(5.0e3), math.log(2.0e4), size=b.tri_count))
    b.tri_materials[:, 0] = ke
    b.color(include_bending=True)
    model, state = _finish(b, device, 0.0, seed, 1.0e4, 0.0, 0.5)
    tag = "_springs" if springs else ""
    info = {"workload": f"cloth_drop_on_sphere_{verts_per_side}x{verts_per_side}{tag}", "dt": FRAME_DT / SUBSTEPS,
            "substeps": SUBSTEPS, "iterations": iterations}
    return model, state, info


def cloth_on_moving_sphere(device="cuda:0", verts_per_side: int = 256, size: float = 2.0, height: float = 1.5,
                           seed: int = 1, iterations: int = 10, sweep: float = 0.3, rise: float = 0.1,
                           spin: float = 1.5):
    """The 256x256 cloth of :func:`cloth_drop_on_sphere` falling onto a sphere carried by body 0, which an
    external rigid solver moves: centre (sweep sin(1.5 t), rise t, 0) and a spin of ``spin`` rad/s about
    y (``info["body_pose"](t)`` gives the [1, 7] pose on the device for a device time scalar ``t``), and
    the static ground.  For ``SolverVBD(integrate_with_external_rigid_solver=True)`` and, optionally, the
    full-surface edge/face contacts."""
    n = verts_per_side - 1
    b = ModelBuilder(up_axis="Y", gravity=-9.81)
    b.add_ground_plane(cfg=ModelBuilder.ShapeConfig(ke=1.0e4, kd=0.0, mu=0.5))
    body = b.add_body(xform=((0.0, 0.0, 0.0), (0.0, 0.0, 0.0, 1.0)), label="mover")
    b.add_shape_sphere(body, radius=0.4, cfg=ModelBuilder.ShapeConfig(ke=1.0e4, kd=0.0, mu=0.18))
    verts, tris = cloth_grid_mesh(n, n, size / n, size / n)
    rot = quat_from_axis_angle((1.0, 0.0, 0.0), math.pi / 2)
    b.add_cloth_mesh(pos=(-0.5 * size, height, -0.5 * size), rot=rot, scale=1.0, vel=(0.0, 0.0, 0.0),
                     vertices=verts, indices=tris, density=0.1, tri_ke=1.0e4, tri_ka=1.0e4, tri_kd=1.0e-5,
                     edge_ke=1.0, edge_kd=0.0)
    rng = np.random.default_rng(seed)
    ke = np.exp(rng.uniform(math.log(5.0e3), math.log(2.0e4), size=b.tri_count))
    b.tri_materials[:, 0] = ke
    b.color(include_bending=True)
    model, state = _finish(b, device, 0.0, seed, 1.0e4, 0.0, 0.5)

    def body_pose(t):
        import torch

        half = 0.5 * spin * t
        zero = torch.zeros_like(t)
        return torch.stack([sweep * torch.sin(1.5 * t), rise * t, zero, zero, torch.sin(half), zero,
                            torch.cos(half)]).reshape(1, 7).to(torch.float32)

    info = {"workload": f"cloth_on_moving_sphere_{verts_per_side}x{verts_per_side}", "dt": FRAME_DT / SUBSTEPS,
            "substeps": SUBSTEPS, "iterations": iterations, "body_pose": body_pose}
    return model, state, info


def cloth_strip_pinned(device="cuda:0", nx: int = 6, ny: int = 4, seed: int = 0, tri_kd: float = 1e-2,
                       edge_kd: float = 1e-3, iterations: int = 3):
    """Tiny cloth (default 6x4 vertices) with membrane and bending damping on, one corner pinned.

    Used by parity tests that want every element type and damping term active.
    """
    b = ModelBuilder(up_axis="Y", gravity=-9.81)
    b.add_ground_plane(cfg=ModelBuilder.ShapeConfig(ke=1.0e4, kd=1.0, mu=0.5))
    b.add_shape_sphere(-1, xform=(np.array([0.1, 0.2, 0.1], np.float32), np.array([0, 0, 0, 1], np.float32)),
                       radius=0.15, cfg=ModelBuilder.ShapeConfig(ke=5.0e3, kd=2.0, mu=0.4))
    verts, tris = cloth_grid_mesh(nx - 1, ny - 1, 0.05, 0.05)
    rot = quat_from_axis_angle((1.0, 0.0, 0.0), math.pi / 2)
    b.add_cloth_mesh(pos=(0.0, 0.35, 0.0), rot=rot, scale=1.0, vel=(0.0, -0.5, 0.0), vertices=verts, indices=tris,
                     density=0.2, tri_ke=2.0e3, tri_ka=3.0e3, tri_kd=tri_kd, edge_ke=0.5, edge_kd=edge_kd,
                     particle_radius=0.02)
    b.particle_mass[0] = 0.0
    b.particle_flags[0] = 0
    b.color(include_bending=True)
    model, state = _finish(b, device, 2.0e-3, seed, 2.0e3, 1.0, 0.6)
    info = {"workload": f"cloth_strip_{nx}x{ny}", "dt": FRAME_DT / SUBSTEPS, "substeps": SUBSTEPS,
            "iterations": iterations}
    return model, state, info


def bead_rope_light(device="cuda:0", beads: int = 12, mass: float = 1.0e-8, spring_ke: float = 1.0e4,
                    height: float = 0.22, iterations: int = 4):
    """A zig-zag rope of nearly massless beads joined by stiff springs, first bead pinned, swinging under
    gravity.  Two alternating colors.

    With ``mass`` 1e-8 kg the inertia term m/h^2 is ~4e-3 against spring stiffness 1e4, so the vertex
    blocks have Frobenius condition numbers far above 1e6: the case the FP64 block solve is for.
    """
    b = ModelBuilder(up_axis="Y", gravity=-9.81)
    for i in range(beads):
        b.add_particle((0.05 * i, height - 0.02 * i, 0.01 * (i % 2)), (0.0, 0.0, 0.0), 0.0 if i == 0 else mass,
                       radius=0.01, flags=0 if i == 0 else int(ParticleFlags.ACTIVE))
    for i in range(beads - 1):
        b.add_spring(i, i + 1, spring_ke, 0.0)
    b.set_coloring([np.arange(0, beads, 2), np.arange(1, beads, 2)])
    model, state = _finish(b, device, 0.0, 0, 1.0e5, 0.0, 0.5)
    info = {"workload": f"bead_rope_{beads}", "dt": FRAME_DT / SUBSTEPS, "substeps": SUBSTEPS,
            "iterations": iterations}
    return model, state, info


def cloth_batched(device="cuda:0", worlds: int = 4096, side_min: int = 16, side_max: int = 64, seed: int = 2,
                  iterations: int = 5, cell: float = 1.0 / 31.0, height: float = 1.0, max_tilt_deg: float = 30.0):
    """Many independent cloth worlds of uneven size solved together (4096 worlds ~ 7.4M vertices by default).

    World w is a square grid of ``s_w ~ U{side_min..side_max}`` vertices per
    side with the cell size of :func:`cloth_hanging` (1/31 m), the same material
    except a per-world ``tri_ke ~ log-uniform[5e3, 2e4]``, 0, 1 or 2 of its two
    top corners pinned (count uniform; a single pinned corner is either one),
    and a tilt uniform in +/-``max_tilt_deg`` about x, centred at ``height``.
    One ground plane (world -1) is shared by every world; soft-contact pairs
    are world-compatible, so each particle pairs with it once.  Worlds are
    colored independently and merged with the reference's plan
    (``builder.py:2924-2939``).  Seeded from ``seed`` in this order: sides,
    ke, corner counts, corner choice, tilts.
    """
    rng = np.random.default_rng(seed)
    sides = rng.integers(side_min, side_max + 1, size=worlds)
    kes = np.exp(rng.uniform(math.log(5.0e3), math.log(2.0e4), size=worlds))
    ncorner = rng.integers(0, 3, size=worlds)
    which = rng.integers(0, 2, size=worlds)
    tilts = np.radians(rng.uniform(-max_tilt_deg, max_tilt_deg, size=worlds))

    templates = {}
    for s in np.unique(sides):
        n = int(s) - 1
        t = ModelBuilder(up_axis="Y", gravity=-9.81)
        verts, tris = cloth_grid_mesh(n, n, cell, cell)
        half = 0.5 * n * cell
        # grid in the XZ plane centred on the origin; each world's xform tilts and lifts it
        t.add_cloth_mesh(pos=(-half, 0.0, -half), rot=quat_from_axis_angle((1.0, 0.0, 0.0), math.pi / 2), scale=1.0,
                         vel=(0.0, 0.0, 0.0), vertices=verts, indices=tris, density=0.1, tri_ke=1.0e4, tri_ka=1.0e4,
                         tri_kd=1.0e-5, edge_ke=1.0, edge_kd=0.0)
        t.color(include_bending=True)
        templates[int(s)] = t

    b = ModelBuilder(up_axis="Y", gravity=-9.81)
    b.add_ground_plane(cfg=ModelBuilder.ShapeConfig(ke=1.0e4, kd=0.0, mu=0.5))
    p_start = np.zeros(worlds + 1, dtype=np.int64)
    t_start = np.zeros(worlds + 1, dtype=np.int64)
    for w in range(worlds):
        t = templates[int(sides[w])]
        p_start[w] = b.particle_count
        t_start[w] = b.tri_count
        q = quat_from_axis_angle((1.0, 0.0, 0.0), float(tilts[w]))
        b.add_world(t, xform=(np.array([0.0, height, 0.0], np.float32), q))
    p_start[worlds] = b.particle_count
    t_start[worlds] = b.tri_count

    tri_mat = b.tri_materials
    tri_mat[:, 0] = np.repeat(kes, np.diff(t_start))
    mass = b.particle_mass
    flags = b.particle_flags
    for w in range(worlds):
        n = int(sides[w]) - 1
        top = [p_start[w] + n * (n + 1), p_start[w] + n * (n + 1) + n]
        pinned = top if ncorner[w] == 2 else ([top[which[w]]] if ncorner[w] == 1 else [])
        for v in pinned:
            mass[v] = 0.0
            flags[v] = int(ParticleFlags.ACTIVE) & ~int(ParticleFlags.ACTIVE)
    model, state = _finish(b, device, 0.0, seed, 1.0e4, 0.0, 0.5)
    info = {"workload": f"cloth_batched_{worlds}w_{side_min}-{side_max}", "dt": FRAME_DT / SUBSTEPS,
            "substeps": SUBSTEPS, "iterations": iterations, "world_particle_start": p_start,
            "world_tri_start": t_start, "sides": sides}
    return model, state, info


class FourierWind:
    """Spatially smooth random wind: an acceleration field that is a sum of Fourier modes.

    ``a(x) = sum_m A_m sin(k_m . x + phi_m)`` with ``A_m ~ N(0, sigma)`` per
    component (N/kg), wave vectors with wavelengths in [0.5, 2] x ``length`` and
    random directions, phases uniform.  ``redraw(frame)`` draws a new set every
    ``period`` frames from one generator seeded with ``seed``.  ``force(q, mass)``
    returns ``mass * a(q)`` ([N,3], the ``particle_f`` the solver reads).
    """

    def __init__(self, device, modes: int = 8, sigma: float = 2.0, length: float = 1.0, period: int = 60,
                 seed: int = 3):
        self.device = torch.device(device)
        self.modes, self.sigma, self.length, self.period = modes, sigma, length, period
        self.rng = np.random.default_rng(seed)
        self.epoch = -1
        self.redraw(0)

    def redraw(self, frame: int) -> None:
        epoch = frame // self.period
        while self.epoch < epoch:
            m = self.modes
            amp = self.rng.normal(0.0, self.sigma, size=(m, 3))
            direction = self.rng.normal(size=(m, 3))
            direction /= np.linalg.norm(direction, axis=1, keepdims=True)
            wavelength = self.length * self.rng.uniform(0.5, 2.0, size=(m, 1))
            k = direction * (2.0 * math.pi / wavelength)
            phase = self.rng.uniform(0.0, 2.0 * math.pi, size=m)
            self.amp = torch.as_tensor(amp, dtype=torch.float32, device=self.device)
            self.k = torch.as_tensor(k, dtype=torch.float32, device=self.device)
            self.phase = torch.as_tensor(phase, dtype=torch.float32, device=self.device)
            self.epoch += 1

    def acceleration(self, q: torch.Tensor) -> torch.Tensor:
        s = torch.sin(q @ self.k.T + self.phase)  # [N, modes]
        return s @ self.amp

    def force(self, q: torch.Tensor, mass: torch.Tensor) -> torch.Tensor:
        return self.acceleration(q) * mass[:, None]


def cloth_wind(device="cuda:0", verts_per_side: int = 128, size: float = 1.0, top: float = 1.5, seed: int = 3):
    """128x128 cloth hanging in the XY plane, pinned along its top edge, driven by :class:`FourierWind`.

    Material of :func:`cloth_hanging` (areal density 0.1, tri_ke = tri_ka = 1e4,
    tri_kd = 1e-5, edge_ke = 1); no shapes.  The wind (8 modes, N(0, 2 N/kg)
    amplitudes, redrawn every 60 frames from ``seed``) is returned in
    ``info["wind"]``; drivers set ``state.particle_f = wind.force(q, mass)``
    before each substep.
    """
    n = verts_per_side - 1
    b = ModelBuilder(up_axis="Y", gravity=-9.81)
    verts, tris = cloth_grid_mesh(n, n, size / n, size / n)
    # grid y runs upward in the XY plane; the row y = n is the top edge
    b.add_cloth_mesh(pos=(-0.5 * size, top - size, 0.0), rot=(0.0, 0.0, 0.0, 1.0), scale=1.0, vel=(0.0, 0.0, 0.0),
                     vertices=verts, indices=tris, density=0.1, tri_ke=1.0e4, tri_ka=1.0e4, tri_kd=1.0e-5,
                     edge_ke=1.0, edge_kd=0.0)
    top_row = n * (n + 1) + np.arange(n + 1)
    b.particle_mass[top_row] = 0.0
    b.particle_flags[top_row] = int(ParticleFlags.ACTIVE) & ~int(ParticleFlags.ACTIVE)
    b.color(include_bending=True)
    model, state = _finish(b, device, 0.0, seed, 1.0e4, 0.0, 0.5)
    wind = FourierWind(device, length=size, seed=seed)
    info = {"workload": f"cloth_wind_{verts_per_side}x{verts_per_side}", "dt": FRAME_DT / SUBSTEPS,
            "substeps": SUBSTEPS, "iterations": 10, "wind": wind}
    return model, state, info


class BatchedFourierWind:
    """One :class:`FourierWind` per world for models of equal-size worlds in contiguous particle ranges.

    World ``w`` draws its modes from ``seed + w`` (world 0 has the wind of the single-world
    scene); ``force(q, mass)`` evaluates every world's field at its own particles.
    """

    def __init__(self, device, worlds: int, per_world: int, seed: int = 3, **kw):
        self.worlds, self.per_world = worlds, per_world
        self.winds = [FourierWind(device, seed=seed + w, **kw) for w in range(worlds)]
        self._stack()

    def _stack(self) -> None:
        self.k_t = torch.stack([w.k.T.contiguous() for w in self.winds])        # [W, 3, m]
        self.phase = torch.stack([w.phase for w in self.winds])[:, None, :]     # [W, 1, m]
        self.amp = torch.stack([w.amp for w in self.winds])                     # [W, m, 3]

    def redraw(self, frame: int) -> None:
        before = [w.epoch for w in self.winds]
        for w in self.winds:
            w.redraw(frame)
        if [w.epoch for w in self.winds] != before:
            self._stack()

    def acceleration(self, q: torch.Tensor) -> torch.Tensor:
        qw = q.reshape(self.worlds, self.per_world, 3).contiguous()
        s = torch.sin(torch.bmm(qw, self.k_t) + self.phase)                     # [W, n, m]
        return torch.bmm(s, self.amp).reshape(-1, 3)

    def force(self, q: torch.Tensor, mass: torch.Tensor) -> torch.Tensor:
        return self.acceleration(q) * mass[:, None]


def cloth_wind_batched(device="cuda:0", worlds: int = 64, verts_per_side: int = 128, size: float = 1.0,
                       top: float = 1.5, seed: int = 3):
    """``worlds`` copies of :func:`cloth_wind` in one model, each world driven by its own wind
    (:class:`BatchedFourierWind`, world ``w`` seeded ``seed + w``); worlds colored independently and
    merged with the reference's plan, one launch per color pass for all of them."""
    n = verts_per_side - 1
    t = ModelBuilder(up_axis="Y", gravity=-9.81)
    verts, tris = cloth_grid_mesh(n, n, size / n, size / n)
    t.add_cloth_mesh(pos=(-0.5 * size, top - size, 0.0), rot=(0.0, 0.0, 0.0, 1.0), scale=1.0, vel=(0.0, 0.0, 0.0),
                     vertices=verts, indices=tris, density=0.1, tri_ke=1.0e4, tri_ka=1.0e4, tri_kd=1.0e-5,
                     edge_ke=1.0, edge_kd=0.0)
    top_row = n * (n + 1) + np.arange(n + 1)
    t.particle_mass[top_row] = 0.0
    t.particle_flags[top_row] = int(ParticleFlags.ACTIVE) & ~int(ParticleFlags.ACTIVE)
    t.color(include_bending=True)
    b = ModelBuilder(up_axis="Y", gravity=-9.81)
    for _ in range(worlds):
        b.add_world(t)
    model, state = _finish(b, device, 0.0, seed, 1.0e4, 0.0, 0.5)
    per = verts_per_side * verts_per_side
    wind = BatchedFourierWind(device, worlds, per, seed=seed, length=size)
    info = {"workload": f"cloth_wind_batched_{worlds}w_{verts_per_side}x{verts_per_side}", "dt": FRAME_DT / SUBSTEPS,
            "substeps": SUBSTEPS, "iterations": 10, "wind": wind}
    return model, state, info


def soft_cubes_batched(device="cuda:0", worlds: int = 512, cells_min: int = 6, cells_max: int = 12,
                       cell: float = 0.05, seed: int = 4, iterations: int = 8, particle_radius: float = 0.01):
    """Many independent tetrahedral cubes (5 tets per hex cell) dropped onto one ground plane.

    World w is a cube of ``n_w ~ U{cells_min..cells_max}`` cells per side
    (``cell`` m each, so (n_w + 1)^3 particles and 5 n_w^3 tets), stable
    Neo-Hookean with Young's modulus log-uniform[1e4, 1e5] Pa and Poisson's
    ratio uniform[0.3, 0.45] (mu = E / (2 (1 + nu)), lambda = E nu / ((1 + nu)
    (1 - 2 nu))), density 1000 kg/m^3, dropped with its centre at a height
    uniform in [0.5, 1.5] m and a uniformly random rotation.  Ground contact
    ke 1e5, mu 0.5.  Seeded from ``seed`` in this order: cells, E, nu, heights,
    rotations.  The surface triangles carry zero material (skipped by the solve).
    """
    rng = np.random.default_rng(seed)
    cells = rng.integers(cells_min, cells_max + 1, size=worlds)
    young = np.exp(rng.uniform(math.log(1.0e4), math.log(1.0e5), size=worlds))
    nu = rng.uniform(0.3, 0.45, size=worlds)
    heights = rng.uniform(0.5, 1.5, size=worlds)
    quats = rng.normal(size=(worlds, 4))
    quats /= np.linalg.norm(quats, axis=1, keepdims=True)

    templates = {}
    for n in np.unique(cells):
        n = int(n)
        t = ModelBuilder(up_axis="Y", gravity=-9.81)
        half = 0.5 * n * cell
        t.add_soft_grid(pos=(-half, -half, -half), rot=(0.0, 0.0, 0.0, 1.0), vel=(0.0, 0.0, 0.0), dim_x=n, dim_y=n,
                        dim_z=n, cell_x=cell, cell_y=cell, cell_z=cell, density=1000.0, k_mu=1.0, k_lambda=1.0,
                        k_damp=0.0, particle_radius=particle_radius)
        t.color(include_bending=False)
        templates[n] = t

    b = ModelBuilder(up_axis="Y", gravity=-9.81)
    b.add_ground_plane(cfg=ModelBuilder.ShapeConfig(ke=1.0e5, kd=0.0, mu=0.5))
    tet_start = np.zeros(worlds + 1, dtype=np.int64)
    p_start = np.zeros(worlds + 1, dtype=np.int64)
    for w in range(worlds):
        p_start[w] = b.particle_count
        tet_start[w] = b.tet_count
        b.add_world(templates[int(cells[w])],
                    xform=(np.array([0.0, heights[w], 0.0], np.float32), quats[w].astype(np.float32)))
    p_start[worlds] = b.particle_count
    tet_start[worlds] = b.tet_count
    mu = young / (2.0 * (1.0 + nu))
    lam = young * nu / ((1.0 + nu) * (1.0 - 2.0 * nu))
    tet_mat = b._tet_mat.array()
    counts = np.diff(tet_start)
    tet_mat[:, 0] = np.repeat(mu, counts)
    tet_mat[:, 1] = np.repeat(lam, counts)
    tet_mat[:, 2] = 0.0
    model, state = _finish(b, device, 0.0, seed, 1.0e5, 0.0, 0.5)
    info = {"workload": f"soft_cubes_batched_{worlds}w_{cells_min}-{cells_max}", "dt": FRAME_DT / SUBSTEPS,
            "substeps": SUBSTEPS, "iterations": iterations, "world_particle_start": p_start,
            "world_tet_start": tet_start, "cells": cells}
    return model, state, info


def soft_cube(device="cuda:0", cells: int = 3, seed: int = 0, iterations: int = 4, k_damp: float = 1e-3):
    """Small tetrahedral cube (5 tets per hex cell) above the ground, for tet-path parity."""
    b = ModelBuilder(up_axis="Y", gravity=-9.81)
    b.add_ground_plane(cfg=ModelBuilder.ShapeConfig(ke=1.0e5, kd=0.0, mu=0.5))
    e, nu = 3.0e4, 0.35
    mu = e / (2 * (1 + nu))
    lam = e * nu / ((1 + nu) * (1 - 2 * nu))
    rot = quat_from_axis_angle((0.3, 1.0, 0.2), 0.4)
    b.add_soft_grid(pos=(0.0, 0.12, 0.0), rot=rot, vel=(0.0, -1.0, 0.0), dim_x=cells, dim_y=cells, dim_z=cells,
                    cell_x=0.05, cell_y=0.05, cell_z=0.05, density=1000.0, k_mu=mu, k_lambda=lam, k_damp=k_damp,
                    particle_radius=0.01)
    b.color(include_bending=False)
    model, state = _finish(b, device, 1.0e-4, seed, 1.0e5, 0.0, 0.5)
    info = {"workload": f"soft_cube_{cells}", "dt": FRAME_DT / SUBSTEPS, "substeps": SUBSTEPS,
            "iterations": iterations}
    return model, state, info


__all__ = ["FRAME_DT", "SUBSTEPS", "FourierWind", "cloth_batched", "cloth_drop_on_sphere", "cloth_hanging",
           "cloth_on_moving_sphere",
           "cloth_strip_pinned", "cloth_wind", "soft_cube", "soft_cubes_batched", "cloth_hanging_self_contact",
           "cloth_layers_on_sphere"]


def cloth_hanging_self_contact(device="cuda:0", width: int = 64, height: int = 32, iterations: int = 10):
    """The reference's hanging-cloth example with VBD self-contact (newton/examples/cloth/example_cloth_hanging.py).

    A width x height cell grid (cell 0.1 m, 0.1 kg per particle, left edge fixed) at z = 4 m (Z up),
    tri_ke = tri_ka = 1e3, tri_kd = 1e2, edge_ke = 10, particle radius 0.05; ground plane with the
    default shape material; soft contact ke 1e2, kd 1e2, mu 1.0; self-contact radius 0.02, margin 0.03;
    60 fps, 10 substeps, 10 iterations.
    """
    b = ModelBuilder(up_axis="Z", gravity=-9.81)
    b.add_ground_plane()
    b.add_cloth_grid(pos=(0.0, 0.0, 4.0), rot=quat_from_axis_angle((0.0, 0.0, 1.0), math.pi * 0.5),
                     vel=(0.0, 0.0, 0.0), dim_x=width, dim_y=height, cell_x=0.1, cell_y=0.1, mass=0.1,
                     fix_left=True, edge_ke=1.0e1, edge_kd=0.0, particle_radius=0.05, tri_ke=1.0e3, tri_ka=1.0e3,
                     tri_kd=1.0e2)
    b.color(include_bending=True)
    model = b.finalize(device)
    model.soft_contact_ke, model.soft_contact_kd, model.soft_contact_mu = 1.0e2, 1.0e2, 1.0
    info = {"workload": f"cloth_hanging_self_contact_{width}x{height}", "dt": FRAME_DT / SUBSTEPS,
            "substeps": SUBSTEPS, "iterations": iterations,
            "solver_kwargs": {"particle_enable_self_contact": True, "particle_self_contact_radius": 0.02,
                              "particle_self_contact_margin": 0.03}}
    return model, model.state(), info


def cloth_layers_on_sphere(device="cuda:0", verts_per_side: int = 128, layers: int = 2, gap: float = 0.008,
                           size: float = 2.0, height: float = 1.0, seed: int = 1, iterations: int = 10):
    """Stacked cloth sheets (the drop-on-sphere material) falling onto a sphere and the ground with
    self-contact: the sheets land on each other and fold over the sphere, so vertex-triangle and
    edge-edge contacts are everywhere.  Self-contact radius 0.4 x the gap, margin 0.75 x the gap."""
    n = verts_per_side - 1
    b = ModelBuilder(up_axis="Y", gravity=-9.81)
    b.add_ground_plane(cfg=ModelBuilder.ShapeConfig(ke=1.0e4, kd=0.0, mu=0.5))
    b.add_shape_sphere(-1, radius=0.4, cfg=ModelBuilder.ShapeConfig(ke=1.0e4, kd=0.0, mu=0.18))
    verts, tris = cloth_grid_mesh(n, n, size / n, size / n)
    rot = quat_from_axis_angle((1.0, 0.0, 0.0), math.pi / 2)
    for k in range(layers):
        b.add_cloth_mesh(pos=(-0.5 * size + 0.3 * gap * k, height + gap * k, -0.5 * size + 0.2 * gap * k), rot=rot,
                         scale=1.0, vel=(0.0, 0.0, 0.0), vertices=verts, indices=tris, density=0.1, tri_ke=1.0e4,
                         tri_ka=1.0e4, tri_kd=1.0e-5, edge_ke=1.0, edge_kd=0.0)
    rng = np.random.default_rng(seed)
    b.tri_materials[:, 0] = np.exp(rng.uniform(math.log(5.0e3), math.log(2.0e4), size=b.tri_count))
    b.color(include_bending=True)
    model, state = _finish(b, device, 0.0, seed, 1.0e4, 0.0, 0.5)
    info = {"workload": f"cloth_layers_on_sphere_{layers}x{verts_per_side}x{verts_per_side}", "dt": FRAME_DT / SUBSTEPS,
            "substeps": SUBSTEPS, "iterations": iterations,
            "solver_kwargs": {"particle_enable_self_contact": True, "particle_self_contact_radius": 0.4 * gap,
                              "particle_self_contact_margin": 0.75 * gap}}
    return model, state, info
