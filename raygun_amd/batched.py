"""Batched scenes (config 3): many independent scenes of uneven size and
resolution, animated per frame, rendered together (rgrt_batch_frame: a
hipGraph per scene on its own stream, or one hipGraph with a branch per
scene), each scene's launch shape picked by the autotuner.

Workload (seeded): triangle counts log-normal, clipped to [10k, 4M]; each
scene a mesh (UV sphere or torus of that size) in 1-6 instances over a ground
plane, config-2 style materials, a resolution from 320x180 to 1920x1080. Per
frame every instance turns about its vertical axis (the TLAS is refitted);
one scene in four deforms its mesh: its vertices are replaced by the next of
a cycle of poses held in device memory (as a GPU skinning pass would leave
them) and its BLAS gets a full LBVH rebuild, both ordered on the scene's
stream with no host wait (rgrt_update_vertices_device, rgrt_rebuild_blas).
No counterpart in the reference, which renders one scene.
"""
import math

import numpy as np

from . import procgen, rgrt
from .hiptime import DeviceBuffer
from .records import INSTANCE, default_material
from .scene import Camera, Model, camera_ubo, pack_models, quat_to_mat3

RESOLUTIONS = [(320, 180), (480, 270), (640, 360), (960, 540), (1280, 720), (1920, 1080)]
POSES = 4  # deformation poses per deforming mesh, cycled frame by frame


def scene_sizes(num_scenes, seed, lo=10_000, hi=4_000_000, sigma=1.2):
    rng = np.random.default_rng(seed)
    mu = 0.5 * (math.log(lo) + math.log(hi))
    return [int(min(hi, max(lo, math.exp(rng.normal(mu, sigma))))) for _ in range(num_scenes)]


class BatchScene:
    """One scene of the batch on its own context."""

    def __init__(self, index, seed, tris, resolution, max_recursions=3, flags=0):
        rng = np.random.default_rng(seed)
        self.index = index
        self.width, self.height = resolution
        self.deforms = index % 4 == 3
        if rng.uniform() < 0.5:
            S = max(8, int(round(math.sqrt(tris))))
            S += S % 2
            self.mesh = procgen.uv_sphere(S, S // 2, 1.0)
        else:
            V = max(4, int(round(math.sqrt(tris / 8))))
            self.mesh = procgen.torus(4 * V, V, 1.0, 0.35)
        ground = procgen.plane(40.0, -1.5)
        self.ctx = rgrt.Context(self.width, self.height, flags)
        models = [Model(self.mesh, np.array([default_material()])), Model(ground, np.array([default_material()]))]
        v, i, _ = pack_models(models)
        self.ctx.upload_geometry(v, i)
        self.vertices, self.indices = v, i
        self.base_vertices = self.mesh.vertices.copy()
        n = int(rng.integers(1, 7))
        mats = [default_material(diffuse=(0.5, 0.5, 0.5), reflectivity=0.3, roughness=0.3, effectId=1)]
        for _ in range(n):
            mats.append(default_material(diffuse=tuple(rng.uniform(0, 1, 3)), roughness=float(rng.uniform(0, 1)),
                                         reflectivity=float(rng.uniform(0.3, 1.0)) if rng.uniform() < 0.4 else 0.0,
                                         transparency=float(rng.uniform(0.5, 0.95)) if rng.uniform() < 0.25 else 0.0,
                                         ior=float(rng.uniform(1.0, 1.8))))
        self.materials = np.array(mats)
        self.ctx.upload_materials(self.materials)
        self.mesh_blas = self.ctx.build_blas(self.mesh.first_vertex, self.mesh.index_offset_bytes, self.mesh.num_faces(),
                                             len(self.mesh.vertices))
        self.ground_blas = self.ctx.build_blas(ground.first_vertex, ground.index_offset_bytes, ground.num_faces(),
                                               len(ground.vertices))
        self.ground = ground
        self.pos = rng.uniform(-4.0, 4.0, size=(n, 3))
        self.pos[:, 1] = rng.uniform(-0.5, 1.5, n)
        self.scale = rng.uniform(0.6, 1.6, n)
        self.speed = rng.uniform(-1.0, 1.0, n)
        cam = Camera(fov_deg=60.0)
        cam.look_at((0.0, 4.0, 14.0), (0.0, 0.0, 0.0))
        self.ubo = camera_ubo(cam, self.width, self.height, max_recursions=max_recursions)
        self.num_instances = n + 1
        self.rays = 0
        self.poses = []
        if self.deforms:
            for k in range(POSES):
                buf = DeviceBuffer(self.base_vertices.nbytes)
                buf.upload(self.pose(2.0 * math.pi * k / POSES))
                self.poses.append(buf)

    def instances(self, t):
        inst = np.zeros(self.num_instances, INSTANCE)
        inst["transform"][0] = [1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0]
        inst["blas"][0] = self.ground_blas
        inst["offsets"][0] = (self.ground.first_vertex, self.ground.index_offset_bytes // 4, 0)
        for k in range(self.num_instances - 1):
            a = self.speed[k] * t
            q = (math.cos(a / 2), 0.0, math.sin(a / 2), 0.0)
            m = np.zeros((3, 4))
            m[:, :3] = quat_to_mat3(q) * self.scale[k]
            m[:, 3] = self.pos[k]
            inst["transform"][k + 1] = m.reshape(-1)
            inst["blas"][k + 1] = self.mesh_blas
            inst["offsets"][k + 1] = (self.mesh.first_vertex, self.mesh.index_offset_bytes // 4, k + 1)
        inst["customIndex"] = np.arange(self.num_instances)
        inst["mask"] = 0xFF
        inst["flags"] = 1
        return inst

    def pose(self, phase):
        """The mesh's vertices under a radial wobble of the given phase."""
        v = self.base_vertices.copy()
        p = v["position"].astype(np.float64)
        w = 1.0 + 0.05 * np.sin(3.0 * p[:, 1] + phase)
        v["position"] = (p * w[:, None]).astype(np.float32)
        return v

    def pose_of_frame(self, frame):
        """The vertices deform(frame) puts in place."""
        return self.pose(2.0 * math.pi * (frame % POSES) / POSES)

    def deform(self, frame):
        """Pose frame % POSES into the vertex buffer and a full BLAS rebuild,
        both queued on the scene's stream."""
        buf = self.poses[frame % POSES]
        self.ctx.update_vertices_device(self.mesh.first_vertex, buf.address, len(self.base_vertices))
        self.ctx.rebuild_blas(self.mesh_blas)

    def close(self):
        self.ctx.close()
        for b in self.poses:
            b.close()
        self.poses = []


class BatchWorkload:
    def __init__(self, num_scenes=32, seed=3, lo=10_000, hi=4_000_000, resolutions=RESOLUTIONS,
                 mode=rgrt.BATCH_GRAPH_PER_SCENE, flags=0):
        rng = np.random.default_rng(seed + 17)
        sizes = scene_sizes(num_scenes, seed, lo, hi)
        self.scenes = [BatchScene(k, seed * 1000 + k, sizes[k], resolutions[int(rng.integers(len(resolutions)))],
                                  flags=flags) for k in range(num_scenes)]
        self.batch = rgrt.Batch([s.ctx for s in self.scenes], mode)
        self.frames_done = 0
        self.info = {"scenes": num_scenes, "triangles": [s.mesh.num_faces() for s in self.scenes],
                     "resolutions": [[s.width, s.height] for s in self.scenes],
                     "pixels": sum(s.width * s.height for s in self.scenes)}

    def descs(self, t):
        first = self.frames_done == 0
        return [(s.ubo, s.instances(t), True, not first) for s in self.scenes]

    def frame_batched(self, t):
        for s in self.scenes:
            if s.deforms and self.frames_done:
                s.deform(self.frames_done)
        self.batch.frame(self.descs(t))
        self.frames_done += 1

    def frame_sequential(self, t):
        for s, (ubo, inst, fxaa, refit) in zip(self.scenes, self.descs(t)):
            if s.deforms and self.frames_done:
                s.deform(self.frames_done)
            s.ctx.frame(ubo, inst, fxaa=fxaa, refit=refit, graph=False)
        self.frames_done += 1

    def synchronize(self):
        for s in self.scenes:
            s.ctx.synchronize()

    def rays(self):
        return sum(s.ctx.stats()["rays"] for s in self.scenes)

    def close(self):
        self.batch.close()
        for s in self.scenes:
            s.close()
