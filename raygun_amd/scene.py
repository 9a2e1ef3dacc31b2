"""Data-compatible subset of Raygun's scene layer, feeding the C ABI.

The reference's Scene / Entity / Transform / Camera / Model / Mesh classes
(raygun/scene.hpp:36-52, raygun/entity.hpp:32-131, raygun/transform.hpp:29-110,
raygun/camera.cpp:34-47, raygun/render/model.hpp:33-42,
raygun/render/mesh.hpp:30-52) are restated here only as far as the
ray-tracing path consumes them:

- ``Transform.to_mat4``          transform.hpp:38-46  (T * R * S); the
  editing helpers ``move`` / ``rotate`` / ``rotate_euler`` /
  ``rotate_around`` / ``look_at`` / ``scale`` of transform.hpp:71-95 and
  ``interpolate`` (:122-131)
- ``Entity.global_transform``    entity.cpp:187-199  (parent * local, the
  parent's global transform cached until the entity or an ancestor moves:
  entity.cpp:180-270)
- ``EntityAnimation``, ``ScaleAnimation``, ``AnimatableEntity``
  entity.hpp:133-175, entity.cpp:300-332; ``Scene.update`` runs them once
  per frame as the main loop does (raygun.cpp:122-126)
- ``Scene.collect_instances``    acceleration_structure.cpp:55-85 (the DFS:
  invisible or zero-volume subtrees are skipped, entities without a model
  are descended into, customIndex = DFS order)
- ``pack_models``                render_system.cpp:192-223, 270-330
  (distinct meshes concatenated in first-use order; each model's materials
  concatenated)
- ``camera_ubo``                 render_system.cpp:235-268, camera.cpp:34-47
"""
import math

import numpy as np

from .records import INSTANCE, MATERIAL, UBO, VERTEX


# ---------------------------------------------------------------------------
# Quaternions / transforms (glm conventions: w, x, y, z; column vectors)
# ---------------------------------------------------------------------------
def quat_to_mat3(q):
    w, x, y, z = (float(c) for c in q)
    return np.array([
        [1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
        [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
        [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)],
    ], dtype=np.float64)


def quat_mul(a, b):
    aw, ax, ay, az = a
    bw, bx, by, bz = b
    return np.array([
        aw * bw - ax * bx - ay * by - az * bz,
        aw * bx + ax * bw + ay * bz - az * by,
        aw * by - ax * bz + ay * bw + az * bx,
        aw * bz + ax * by - ay * bx + az * bw,
    ])


def quat_look_at(direction, up=(0.0, 1.0, 0.0)):
    """glm::quatLookAt (right-handed): the rotation whose -Z axis is `direction`."""
    d = np.asarray(direction, np.float64)
    d = d / np.linalg.norm(d)
    right = np.cross(np.asarray(up, np.float64), -d)
    right /= np.linalg.norm(right)
    upv = np.cross(-d, right)
    m = np.stack([right, upv, -d], axis=1)  # columns
    return mat3_to_quat(m)


def mat3_to_quat(m):
    tr = m[0, 0] + m[1, 1] + m[2, 2]
    if tr > 0:
        s = math.sqrt(tr + 1.0) * 2
        return np.array([0.25 * s, (m[2, 1] - m[1, 2]) / s, (m[0, 2] - m[2, 0]) / s, (m[1, 0] - m[0, 1]) / s])
    if m[0, 0] > m[1, 1] and m[0, 0] > m[2, 2]:
        s = math.sqrt(1.0 + m[0, 0] - m[1, 1] - m[2, 2]) * 2
        return np.array([(m[2, 1] - m[1, 2]) / s, 0.25 * s, (m[0, 1] + m[1, 0]) / s, (m[0, 2] + m[2, 0]) / s])
    if m[1, 1] > m[2, 2]:
        s = math.sqrt(1.0 + m[1, 1] - m[0, 0] - m[2, 2]) * 2
        return np.array([(m[0, 2] - m[2, 0]) / s, (m[0, 1] + m[1, 0]) / s, 0.25 * s, (m[1, 2] + m[2, 1]) / s])
    s = math.sqrt(1.0 + m[2, 2] - m[0, 0] - m[1, 1]) * 2
    return np.array([(m[1, 0] - m[0, 1]) / s, (m[0, 2] + m[2, 0]) / s, (m[1, 2] + m[2, 1]) / s, 0.25 * s])


class Transform:
    """raygun/transform.hpp:29-110 (position, rotation quaternion, scaling)."""

    def __init__(self, position=(0, 0, 0), rotation=(1, 0, 0, 0), scaling=(1, 1, 1)):
        self.position = np.asarray(position, np.float64)
        self.rotation = np.asarray(rotation, np.float64)
        self.scaling = np.asarray(scaling, np.float64)

    def to_mat4(self):
        """toMat4 = translate * toMat4(rotation) * scale (transform.hpp:38-46)."""
        m = np.eye(4)
        m[:3, :3] = quat_to_mat3(self.rotation) * self.scaling[None, :]
        m[:3, 3] = self.position
        return m

    def is_zero_volume(self):
        return float(self.scaling[0] * self.scaling[1] * self.scaling[2]) == 0.0

    def copy(self):
        return Transform(self.position.copy(), self.rotation.copy(), self.scaling.copy())

    def is_identity(self):
        return bool(np.all(self.position == 0) and np.all(self.rotation == (1, 0, 0, 0)) and np.all(self.scaling == 1))

    def inverse(self):
        """transform.hpp:55-62 (component-wise, as the reference defines it)."""
        w, x, y, z = self.rotation
        n = w * w + x * x + y * y + z * z
        return Transform(-self.position, np.array([w, -x, -y, -z]) / n, 1.0 / self.scaling)

    def up(self):
        return quat_to_mat3(self.rotation) @ np.array([0.0, 1.0, 0.0])

    def right(self):
        return quat_to_mat3(self.rotation) @ np.array([1.0, 0.0, 0.0])

    def forward(self):
        return quat_to_mat3(self.rotation) @ np.array([0.0, 0.0, -1.0])

    def move(self, translation):
        self.position = self.position + np.asarray(translation, np.float64)

    def rotate(self, angle, axis):
        """glm::rotate(rotation, angle, axis): rotation * angleAxis (the axis
        normalised when its length is off by more than 0.001)."""
        a = np.asarray(axis, np.float64)
        ln = float(np.linalg.norm(a))
        if abs(ln - 1.0) > 0.001:
            a = a / ln
        sn = math.sin(angle * 0.5)
        self.rotation = quat_mul(self.rotation, (math.cos(angle * 0.5), a[0] * sn, a[1] * sn, a[2] * sn))

    def rotate_euler(self, angles):
        """rotation = quat(eulerAngles) * rotation (transform.hpp:73)."""
        self.rotation = quat_mul(quat_from_euler(angles), self.rotation)

    def rotate_around(self, pivot, angles):
        """transform.hpp:75-80: move(-pivot), rotate(angles), move(pivot). The
        rotation does not touch the position, so only the orientation
        changes, as in the reference."""
        pv = np.asarray(pivot, np.float64)
        self.move(-pv)
        self.rotate_euler(angles)
        self.move(pv)

    def look_at(self, target):
        d = np.asarray(target, np.float64) - self.position
        self.rotation = quat_look_at(d / np.linalg.norm(d))

    def scale(self, factors):
        self.scaling = self.scaling * np.asarray(factors, np.float64)

    def __mul__(self, y):
        """operator* (transform.hpp:113-120)."""
        r = Transform()
        r.position = quat_to_mat3(self.rotation) @ (self.scaling * y.position) + self.position
        r.rotation = quat_mul(self.rotation, y.rotation)
        r.scaling = self.scaling * y.scaling
        return r


def quat_from_euler(angles):
    """glm quat(vec3 eulerAngles) (pitch x, yaw y, roll z)."""
    c = np.cos(np.asarray(angles, np.float64) * 0.5)
    sn = np.sin(np.asarray(angles, np.float64) * 0.5)
    return np.array([
        c[0] * c[1] * c[2] + sn[0] * sn[1] * sn[2],
        sn[0] * c[1] * c[2] - c[0] * sn[1] * sn[2],
        c[0] * sn[1] * c[2] + sn[0] * c[1] * sn[2],
        c[0] * c[1] * sn[2] - sn[0] * sn[1] * c[2],
    ])


def interpolate(x, y, factor):
    """transform.hpp:122-131: component-wise lerp, factor clamped to [0, 1]
    (the quaternion too, without renormalisation, as glm::lerp does)."""
    f = min(max(float(factor), 0.0), 1.0)
    return Transform(x.position * (1 - f) + y.position * f, x.rotation * (1 - f) + y.rotation * f,
                     x.scaling * (1 - f) + y.scaling * f)


class Mesh:
    """raygun/render/mesh.hpp:30-52: vertices (VERTEX records) + mesh-local indices."""

    def __init__(self, vertices, indices, name=""):
        self.vertices = np.ascontiguousarray(vertices, VERTEX)
        self.indices = np.ascontiguousarray(indices, np.uint32)
        self.name = name
        self.first_vertex = 0
        self.index_offset_bytes = 0

    def num_faces(self):
        return len(self.indices) // 3


class Model:
    """raygun/render/model.hpp:33-42: a mesh and its material list."""

    def __init__(self, mesh, materials):
        self.mesh = mesh
        self.materials = np.atleast_1d(np.asarray(materials, MATERIAL))
        self.material_offset = 0
        self.blas = None


class Entity:
    """raygun/entity.hpp:32-131 (transform, optional model, children, visibility).

    ``transform`` is the local transform; assigning it (or calling one of the
    editing methods) drops the cached parent transform of every descendant,
    as Entity::setTransform does (entity.cpp:180-185). Changing the fields of
    ``transform`` in place bypasses that: use ``set_transform``."""

    def __init__(self, name="", transform=None, model=None, visible=True):
        self.name = name
        self._transform = transform or Transform()
        self._cached_parent = None
        self.model = model
        self.visible = visible
        self.children = []
        self.parent = None

    @property
    def transform(self):
        return self._transform

    @transform.setter
    def transform(self, t):
        self.set_transform(t)

    def set_transform(self, t):
        self._invalidate_children()
        self._transform = t

    def add_child(self, e):
        e.set_parent(self)
        self.children.append(e)
        return e

    def set_parent(self, parent):
        self._invalidate()
        self.parent = parent

    def _invalidate(self):
        self._cached_parent = None
        self._invalidate_children()

    def _invalidate_children(self):
        for c in self.children:
            c._invalidate()

    def _edit(self, f):
        self._invalidate_children()
        f(self._transform)

    def move(self, translation):
        self._edit(lambda t: t.move(translation))

    def move_to(self, position):
        self._edit(lambda t: setattr(t, "position", np.asarray(position, np.float64)))

    def rotate(self, angle, axis):
        self._edit(lambda t: t.rotate(angle, axis))

    def rotate_euler(self, angles):
        self._edit(lambda t: t.rotate_euler(angles))

    def rotate_around(self, pivot, angles):
        self._edit(lambda t: t.rotate_around(pivot, angles))

    def look_at_target(self, target):
        self._edit(lambda t: t.look_at(target))

    def scale(self, factors):
        self._edit(lambda t: t.scale(factors))

    def parent_transform(self):
        """entity.cpp:187-194: the parent's global transform, cached."""
        if self._cached_parent is None:
            self._cached_parent = self.parent.global_transform() if self.parent is not None else Transform()
        return self._cached_parent

    def global_transform(self):
        """entity.cpp:196-199: parentTransform() * m_transform."""
        return self.parent_transform() * self._transform

    def for_each_entity(self, f):
        """entity.hpp:67-84: pre-order; f returns False to skip the subtree."""
        if f(self) is False:
            return
        for c in self.children:
            c.for_each_entity(f)


class EntityAnimation:
    """entity.hpp:133-141, entity.cpp:300-304: accumulates time and runs."""

    def __init__(self):
        self.animation_time = 0.0

    def update(self, delta_time, target):
        self.animation_time += delta_time
        return self.run_animation(target)

    def run_animation(self, target):
        raise NotImplementedError


class ScaleAnimation(EntityAnimation):
    """entity.cpp:306-320: scaling = mix(start, end, clamp(t / duration, 0, 1));
    runs while t <= duration."""

    def __init__(self, duration, start_scale, end_scale):
        super().__init__()
        self.duration = float(duration)
        self.start_scale = np.asarray(start_scale, np.float64)
        self.end_scale = np.asarray(end_scale, np.float64)

    def run_animation(self, target):
        t = self.animation_time / self.duration
        f = min(max(t, 0.0), 1.0)
        tr = target.transform.copy()
        tr.scaling = self.start_scale * (1 - f) + self.end_scale * f
        target.set_transform(tr)
        return t <= 1.0


class AnimatableEntity(Entity):
    """entity.hpp:155-175, entity.cpp:322-332."""

    def __init__(self, name="", transform=None, model=None, visible=True):
        super().__init__(name, transform, model, visible)
        self.animation = None
        self.animation_finisher = None

    def set_animation(self, animation):
        self.animation = animation

    def set_animation_finisher(self, f):
        self.animation_finisher = f

    def update(self, delta_time):
        if self.animation is not None:
            if not self.animation.update(delta_time, self):
                self.animation = None
                if self.animation_finisher is not None:
                    self.animation_finisher()
                self.animation_finisher = None


class Scene:
    """raygun/scene.hpp:36-52: a camera and the entity tree."""

    def __init__(self):
        self.root = Entity("root")
        self.camera = Camera()

    def update(self, delta_time):
        """The per-frame animation step of the main loop (raygun.cpp:122-126)."""
        self.root.for_each_entity(lambda e: e.update(delta_time) if isinstance(e, AnimatableEntity) else None)

    def models(self):
        out = []
        self.root.for_each_entity(lambda e: out.append(e.model) if e.model is not None and e.model not in out else None)
        return out

    def collect_instances(self):
        """TopLevelAS(cmd, scene) host loop (acceleration_structure.cpp:55-85)."""
        recs = []

        def visit(e):
            if not e.visible:
                return False
            if e.transform.is_zero_volume():
                return False
            if e.model is None:
                return True
            m4 = e.global_transform().to_mat4()
            r = np.zeros((), INSTANCE)
            r["transform"] = m4[:3, :].reshape(12).astype(np.float32)
            r["customIndex"] = len(recs)
            r["mask"] = 0xFF
            r["flags"] = 1
            r["blas"] = e.model.blas
            r["offsets"] = (e.model.mesh.first_vertex, e.model.mesh.index_offset_bytes // 4, e.model.material_offset)
            recs.append(r)
            return True

        self.root.for_each_entity(visit)
        return np.array(recs, INSTANCE) if recs else np.zeros(0, INSTANCE)


def pack_models(models):
    """RenderSystem::setupModelBuffers (render_system.cpp:192-223, 270-330).

    Returns (vertices, indices, materials) and sets each mesh's
    first_vertex / index_offset_bytes and each model's material_offset.
    """
    meshes = []
    for m in models:
        if all(m.mesh is not x for x in meshes):
            meshes.append(m.mesh)
    vo = io = 0
    for mesh in meshes:
        mesh.first_vertex = vo
        mesh.index_offset_bytes = io * 4
        vo += len(mesh.vertices)
        io += len(mesh.indices)
    vertices = np.concatenate([m.vertices for m in meshes]) if meshes else np.zeros(0, VERTEX)
    indices = np.concatenate([m.indices for m in meshes]) if meshes else np.zeros(0, np.uint32)
    mo = 0
    mats = []
    for m in models:
        m.material_offset = mo
        mats.append(m.materials)
        mo += len(m.materials)
    materials = np.concatenate(mats) if mats else np.zeros(0, MATERIAL)
    return vertices, indices, materials


# ---------------------------------------------------------------------------
# Camera and frame constants
# ---------------------------------------------------------------------------
class Camera(Entity):
    """raygun/camera.hpp:29-48; FOV is a parameter (reference: 45 degrees)."""

    def __init__(self, fov_deg=45.0, near=0.1, far=100.0):
        super().__init__("camera")
        self.fov_deg = fov_deg
        self.near = near
        self.far = far

    def look_at(self, eye, target):
        t = self.transform.copy()
        t.position = np.asarray(eye, np.float64)
        t.rotation = quat_look_at(np.asarray(target, np.float64) - t.position)
        self.set_transform(t)

    def projection(self, width, height):
        """glm::perspectiveRH_ZO (GLM_FORCE_DEPTH_ZERO_TO_ONE) with [1][1] *= -1 (camera.cpp:34-47).

        Returned as a row-major numpy matrix (m[row, col])."""
        aspect = float(np.float32(width) / np.float32(height))
        if not math.isfinite(aspect):
            aspect = 16.0 / 9.0
        f = float(np.float32(math.radians(self.fov_deg)))
        th = math.tan(f / 2.0)
        p = np.zeros((4, 4))
        p[0, 0] = 1.0 / (aspect * th)
        p[1, 1] = 1.0 / th
        p[2, 2] = self.far / (self.near - self.far)
        p[3, 2] = -1.0
        p[2, 3] = -(self.far * self.near) / (self.far - self.near)
        p[1, 1] *= -1.0
        return p


def camera_ubo(camera, width, height, light_dir=(0.4, -0.6, -0.8), num_samples=1, max_recursions=5, fade=(0, 0, 0, 0)):
    """UBO as RenderSystem::resetUniformBuffer + updateUniformBuffer fill it
    (render_system.cpp:235-268). Matrices are stored column-major (glm)."""
    u = np.zeros((), UBO)
    view_inv = camera.global_transform().to_mat4()
    proj_inv = np.linalg.inv(camera.projection(width, height))
    u["viewInverse"] = view_inv.T.reshape(16).astype(np.float32)
    u["projInverse"] = proj_inv.T.reshape(16).astype(np.float32)
    u["clearColor"] = (0.2, 0.2, 0.2)
    u["numSamples"] = num_samples
    ld = np.asarray(light_dir, np.float64)
    u["lightDir"] = (ld / np.linalg.norm(ld)).astype(np.float32)
    u["maxRecursions"] = max_recursions
    u["time"] = 0.0
    u["showAlpha"] = 0
    u["fadeColor"] = fade
    return u
