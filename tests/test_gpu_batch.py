"""Batched scenes: every scene of one batched launch per stage is bitwise
equal to the same scene run alone through FluidSimulator (and to the
oracle), for uneven sizes, per-scene parameters, obstacle maps and splats,
under every Jacobi tile shape, eagerly and under HIP-graph replay."""
import numpy as np
import pytest

from natrix_amd.batch import BatchedScenes, random_rect_obstacles
from oracle import fluid as F
from tests.helpers import assert_bitwise, f32, make_sim, random_velocity

pytestmark = pytest.mark.gpu

SIZES = [(40, 40), (64, 64), (96, 96), (33, 33), (128, 128), (50, 70), (17, 129)]


def _setup(seed=7):
    rng = np.random.default_rng(seed)
    scenes = []
    for i, (w, h) in enumerate(SIZES):
        prm = dict(vorticity=float(rng.uniform(0, 0.6)),
                   viscosity=[0.0, 0.2][i % 2],
                   bc_mode=i % 3, speed=float(rng.uniform(100, 500)),
                   dissipation=float(rng.uniform(0.97, 1.0)))
        u = random_velocity(rng, h, w, 0.3)
        solid = random_rect_obstacles(rng, w, h, rng.uniform(0, 0.15))
        splats = [(i, float(rng.uniform(0.2, 0.8)), float(rng.uniform(0.2, 0.8)),
                   float(rng.normal()), float(rng.normal()), 0.02 * w + 1.0, 1.0)
                  for _ in range(int(rng.integers(1, 4)))]
        scenes.append((prm, u, solid, splats))
    return scenes


def _batch(scenes, iters, tile=None, graph=False):
    b = BatchedScenes([w for w, _ in SIZES], [h for _, h in SIZES], dye=True)
    b.set_params(iterations=iters, **({"jacobi_tile": tile} if tile else {}))
    splats = []
    for i, (prm, u, solid, sp) in enumerate(scenes):
        b.set_scene_params(i, **prm)
        b.upload(i, "velocity", u)
        b.set_obstacle_map(i, solid)
        splats += sp
    b.set_step_splats(splats)
    b.use_graph = graph
    return b


@pytest.mark.parametrize("tile", [None, (6, 4, 8, 2), (3, 8, 8, 1), (8, 2, 16, 2)])
def test_batch_matches_single_scenes(gpu, tile):
    scenes = _setup()
    iters, steps, dt = 13, 4, np.float32(1 / 60)
    b = _batch(scenes, iters, tile)
    for _ in range(steps):
        b.update(dt)
    for i, (prm, u, solid, sp) in enumerate(scenes):
        w, h = SIZES[i]
        sim = make_sim(h, w, dye=True, bc_mode=prm["bc_mode"])
        for k in ("vorticity", "viscosity", "speed", "dissipation"):
            setattr(sim, k, prm[k])
        sim.iterations = iters
        if tile:
            sim.jacobi_tile = tile
        sim.upload("velocity", u)
        sim.set_obstacle_map(solid)
        for _ in range(steps):
            sim.add_splats([s[1:] for s in sp])
            sim.update(dt)
        for f in ("velocity", "pressure", "dye", "mask"):
            assert_bitwise(b.download(i, f), sim.download(f), f"scene {i} {f}")
        sim.destroy()
    b.destroy()


def test_batch_matches_oracle(gpu):
    scenes = _setup(11)
    iters, steps, dt = 9, 3, np.float32(1 / 60)
    b = _batch(scenes, iters)
    for _ in range(steps):
        b.update(dt)
    for i in (0, 5, 6):
        prm, u, solid, sp = scenes[i]
        w, h = SIZES[i]
        op = F.Params(iterations=iters, **{k: prm[k] for k in
                                           ("vorticity", "viscosity", "bc_mode",
                                            "speed", "dissipation")})
        q = np.zeros((h, w), f32)
        p = np.zeros((h, w), f32)
        obs = np.zeros((h, w, 2), f32)
        uu = u
        for _ in range(steps):
            uu, q = F.gaussian_splat(uu, q, [tuple(np.float32(v) for v in s[1:]) for s in sp])
            uu, p, obs, q = F.step(uu, p, obs, dt, op, dye=q, persistent=solid)
        assert_bitwise(b.download(i, "velocity"), uu, f"oracle scene {i}")
        assert_bitwise(b.download(i, "dye"), q, f"oracle dye {i}")
    b.destroy()


def test_batch_graph_replay_matches_eager(gpu):
    scenes = _setup(3)
    a = _batch(scenes, 11, graph=False)
    g = _batch(scenes, 11, graph=True)
    for step in range(5):
        a.update(1 / 60)
        g.update(1 / 60)
        for i in (0, 3, 4):
            assert_bitwise(g.download(i, "velocity"), a.download(i, "velocity"),
                           f"graph step {step} scene {i}")
    a.destroy()
    g.destroy()


def test_batch_per_step_obstacles_equal_single_grids(gpu):
    """Circle and triangle obstacles added to single scenes of a batch before
    each step: every scene bitwise equal to a single grid given the same
    obstacles (static and dynamic triangles, the end-of-step clear)."""
    from natrix_amd import FluidSimulator
    from natrix_amd.batch import BatchedScenes
    sides = [(64, 48), (96, 96), (50, 70)]
    b = BatchedScenes([w for w, h in sides], [h for w, h in sides], dye=True)
    b.set_params(iterations=12, vorticity=0.3, viscosity=0.0, bc_mode="inflow")
    sims = []
    for w, h in sides:
        s = FluidSimulator(w, h, dye=True, bc_mode="inflow")
        s.vorticity, s.viscosity, s.iterations = 0.3, 0.0, 12
        sims.append(s)
    for step in range(4):
        c = ((0.3 + 0.05 * step, 0.5), 6.0)
        t = ((0.5, 0.2), (0.8, 0.5 + 0.02 * step), (0.55, 0.8))
        b.add_circle_obstacle(1, *c)
        sims[1].add_circle_obstacle(*c)
        b.add_triangle_obstacle(2, *t, static=step % 2 == 1)
        sims[2].add_triangle_obstacle(*t, static=step % 2 == 1)
        b.update(1 / 60)
        for s in sims:
            s.update(1 / 60)
    for i, s in enumerate(sims):
        for f in ("velocity", "pressure", "dye", "mask"):
            assert np.array_equal(b.download(i, f), s.download(f)), (i, f)
        s.destroy()
    b.destroy()
