"""Host-side mirror of Raygun's renderer classes over the C ABI.

Same method names and call order as the reference's host code:

- ``RenderSystem.setupModelBuffers``  raygun/render/render_system.cpp:192-223
- ``Raytracer.setupBottomLevelAS``     raygun/render/raytracer.cpp:54-74
- ``Raytracer.setupTopLevelAS``        raygun/render/raytracer.cpp:76-85
  (optionally a refit when only transforms changed, ``refit_tlas``)
- ``RenderSystem.updateUniformBuffer`` raygun/render/render_system.cpp:246-268
- ``Raytracer.doRaytracing``           raygun/render/raytracer.cpp:87-147
- ``RenderSystem.render``              raygun/render/render_system.cpp:88-162
  (minus swapchain blit, ImGui and present, which are out of scope)

One BLAS is built per distinct mesh (the reference builds one per model;
models sharing a mesh get identical trees, so this only saves work).
"""
import numpy as np

from . import rgrt
from .scene import camera_ubo, pack_models


class Raytracer:
    def __init__(self, ctx):
        self.ctx = ctx
        self.use_fxaa = True  # m_useFXAA (raytracer.hpp:80)
        # The reference rebuilds the TLAS every frame. With refit_tlas set, a
        # frame whose instances keep their count and BLASes (only transforms
        # changed) refits the previous tree instead.
        self.refit_tlas = False
        self._last_blas = None

    def setupBottomLevelAS(self, models):
        built = {}
        for model in models:
            mesh = model.mesh
            key = id(mesh)
            if key not in built:
                if getattr(mesh, "blas", None) is None:
                    mesh.blas = self.ctx.build_blas(mesh.first_vertex, mesh.index_offset_bytes, mesh.num_faces(),
                                                    len(mesh.vertices))
                built[key] = mesh.blas
            model.blas = built[key]

    def setupTopLevelAS(self, scene):
        inst = scene.collect_instances()
        blas = inst["blas"].copy()
        if self.refit_tlas and self._last_blas is not None and np.array_equal(blas, self._last_blas):
            self.ctx.refit_tlas(inst)
        else:
            self.ctx.build_tlas(inst)
        self._last_blas = blas

    def updateRenderTarget(self, ubo):
        self.ctx.set_frame(ubo)

    def doRaytracing(self):
        self.ctx.render(self.use_fxaa)
        return rgrt.IMAGE_FINAL

    def doFrame(self, scene, ubo, graph=True):
        """setupTopLevelAS + doRaytracing with the frame constants as one
        rgrt_frame call (one hipGraph launch when graph is set)."""
        inst = scene.collect_instances()
        blas = inst["blas"].copy()
        refit = self.refit_tlas and self._last_blas is not None and np.array_equal(blas, self._last_blas)
        self.ctx.frame(ubo, inst, fxaa=self.use_fxaa, refit=refit, graph=graph)
        self._last_blas = blas
        return rgrt.IMAGE_FINAL


# Launch shape of the frame when one frame is in flight: traced in three
# strips side by side (frameBands), 5 wavefront passes, 4/8/2 trace/shade/
# drain blocks per CU (measured on the 1080p 2M-triangle scene). With frames
# in flight the contexts already run side by side and keep the library's
# one-band default.
ONE_FRAME_TUNING = {"frameBands": 3, "wavefrontPasses": 5, "traceBlocksPerCU": 4, "shadeBlocksPerCU": 8,
                    "drainBlocksPerCU": 2}


class RenderSystem:
    def __init__(self, width, height, flags=0, device=0):
        self.ctx = rgrt.Context(width, height, flags, device)
        self.raytracer = Raytracer(self.ctx)
        self.width, self.height = width, height
        self._extra = []  # (context, Raytracer) per frame in flight beyond the first
        self._frame = 0
        self._last = self.ctx
        self._default_tuning = self.ctx.tuning()
        self._apply_tuning()

    def _apply_tuning(self):
        """Sets the launch shape on every context: ONE_FRAME_TUNING with one
        frame in flight, the library default with more (no counterpart in
        the reference)."""
        t = dict(self._default_tuning)
        if not self._extra:
            t.update(ONE_FRAME_TUNING)
        for c in [self.ctx] + [c for c, _ in self._extra]:
            c.set_tuning(**t)

    def set_frames_in_flight(self, n):
        """n - 1 more contexts sharing the scene (rgrt_create_shared):
        render_frame then rotates over the n contexts, so a frame starts
        while the previous ones still run; read() reads the latest frame.
        No counterpart in the reference, which renders one frame at a time."""
        if n < 1:
            raise ValueError("frames in flight must be at least 1")
        for c, _ in self._extra:
            c.close()
        self._extra = []
        for _ in range(n - 1):
            c = self.ctx.shared()
            self._extra.append((c, Raytracer(c)))
        self._apply_tuning()
        self._frame = 0
        self._last = self.ctx

    def setupModelBuffers(self, scene):
        models = scene.models()
        for m in models:
            m.mesh.blas = None
        for rt in [self.raytracer] + [r for _, r in self._extra]:
            rt._last_blas = None  # new geometry drops every sharing context's TLAS
        vertices, indices, materials = pack_models(models)
        self.ctx.upload_geometry(vertices, indices)
        self.ctx.upload_materials(materials)
        return models

    def load_scene(self, scene):
        """Raygun::finalizeLoadScene's render part (raygun.cpp:276-295)."""
        models = self.setupModelBuffers(scene)
        self.raytracer.setupBottomLevelAS(models)

    def updateUniformBuffer(self, camera, max_recursions=5, num_samples=1, light_dir=(0.4, -0.6, -0.8), fade=(0, 0, 0, 0)):
        ubo = camera_ubo(camera, self.width, self.height, light_dir=light_dir, num_samples=num_samples,
                         max_recursions=max_recursions, fade=fade)
        self.raytracer.updateRenderTarget(ubo)
        self._frame_settings = dict(light_dir=light_dir, num_samples=num_samples, max_recursions=max_recursions, fade=fade)
        return ubo

    def render_frame(self, scene, graph=True):
        """The whole frame (frame constants from the scene camera with the
        settings of the last updateUniformBuffer, TLAS, trace, post) as one
        rgrt_frame call."""
        ubo = camera_ubo(scene.camera, self.width, self.height, **getattr(self, "_frame_settings", {}))
        k = self._frame % (1 + len(self._extra))
        self._frame += 1
        ctx, rt = (self.ctx, self.raytracer) if k == 0 else self._extra[k - 1]
        rt.use_fxaa, rt.refit_tlas = self.raytracer.use_fxaa, self.raytracer.refit_tlas
        self._last = ctx
        return rt.doFrame(scene, ubo, graph)

    def render(self, scene, ubo=None):
        self.raytracer.setupTopLevelAS(scene)
        if ubo is not None:
            self.raytracer.updateRenderTarget(ubo)
        self._last = self.ctx
        return self.raytracer.doRaytracing()

    def read(self, which=rgrt.IMAGE_FINAL):
        return self._last.read_image(which)

    def close(self):
        for c, _ in self._extra:
            c.close()
        self._extra = []
        self.ctx.close()


def half_to_float(img16):
    return np.asarray(img16, np.uint16).view(np.float16).astype(np.float32)
