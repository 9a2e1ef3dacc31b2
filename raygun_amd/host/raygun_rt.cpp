// raygun_rt.cpp — C++ host side of Raygun's ray-tracing path over the rgrt C ABI.
// Class-by-class sources are listed in raygun_rt.hpp.
#include "raygun_rt.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>

namespace raygun {

namespace {

void ok(rgrt_ctx* ctx, rgrt_status s, const char* what)
{
    if(s == RGRT_OK) return;
    const char* msg = ctx ? rgrt_last_error(ctx) : nullptr;
    throw Error(s, std::string(what) + ": " + (msg && *msg ? msg : "rgrt error " + std::to_string(s)));
}

vec3 cross(const vec3& a, const vec3& b)
{
    return {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]};
}

vec3 normalized(const vec3& a)
{
    const double l = std::sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]);
    return {a[0] / l, a[1] / l, a[2] / l};
}

// rotation matrix of a unit quaternion (glm::toMat3), m[row][col]
std::array<std::array<double, 3>, 3> quatToMat3(const quat& q)
{
    const double w = q[0], x = q[1], y = q[2], z = q[3];
    return {{{1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)},
             {2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)},
             {2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)}}};
}

quat quatMul(const quat& a, const quat& b)
{
    return {a[0] * b[0] - a[1] * b[1] - a[2] * b[2] - a[3] * b[3], a[0] * b[1] + a[1] * b[0] + a[2] * b[3] - a[3] * b[2],
            a[0] * b[2] - a[1] * b[3] + a[2] * b[0] + a[3] * b[1], a[0] * b[3] + a[1] * b[2] - a[2] * b[1] + a[3] * b[0]};
}

quat mat3ToQuat(const std::array<std::array<double, 3>, 3>& m)
{
    const double tr = m[0][0] + m[1][1] + m[2][2];
    if(tr > 0) {
        const double s = std::sqrt(tr + 1.0) * 2;
        return {0.25 * s, (m[2][1] - m[1][2]) / s, (m[0][2] - m[2][0]) / s, (m[1][0] - m[0][1]) / s};
    }
    if(m[0][0] > m[1][1] && m[0][0] > m[2][2]) {
        const double s = std::sqrt(1.0 + m[0][0] - m[1][1] - m[2][2]) * 2;
        return {(m[2][1] - m[1][2]) / s, 0.25 * s, (m[0][1] + m[1][0]) / s, (m[0][2] + m[2][0]) / s};
    }
    if(m[1][1] > m[2][2]) {
        const double s = std::sqrt(1.0 + m[1][1] - m[0][0] - m[2][2]) * 2;
        return {(m[0][2] - m[2][0]) / s, (m[0][1] + m[1][0]) / s, 0.25 * s, (m[1][2] + m[2][1]) / s};
    }
    const double s = std::sqrt(1.0 + m[2][2] - m[0][0] - m[1][1]) * 2;
    return {(m[1][0] - m[0][1]) / s, (m[0][2] + m[2][0]) / s, (m[1][2] + m[2][1]) / s, 0.25 * s};
}

} // namespace

mat4 identity4()
{
    mat4 m{};
    for(int i = 0; i < 4; ++i) m[i][i] = 1.0;
    return m;
}

mat4 mul(const mat4& a, const mat4& b)
{
    mat4 r{};
    for(int i = 0; i < 4; ++i)
        for(int j = 0; j < 4; ++j) {
            double s = 0.0;
            for(int k = 0; k < 4; ++k) s += a[i][k] * b[k][j];
            r[i][j] = s;
        }
    return r;
}

// Gauss-Jordan elimination with partial pivoting, in double.
mat4 inverse(const mat4& m)
{
    double a[4][8];
    for(int i = 0; i < 4; ++i)
        for(int j = 0; j < 4; ++j) {
            a[i][j] = m[i][j];
            a[i][4 + j] = i == j ? 1.0 : 0.0;
        }
    for(int c = 0; c < 4; ++c) {
        int p = c;
        for(int r = c + 1; r < 4; ++r)
            if(std::fabs(a[r][c]) > std::fabs(a[p][c])) p = r;
        if(a[p][c] == 0.0) throw Error(RGRT_ERR_INVALID_ARG, "inverse: singular matrix");
        if(p != c)
            for(int j = 0; j < 8; ++j) std::swap(a[p][j], a[c][j]);
        const double d = a[c][c];
        for(int j = 0; j < 8; ++j) a[c][j] /= d;
        for(int r = 0; r < 4; ++r) {
            if(r == c) continue;
            const double f = a[r][c];
            if(f == 0.0) continue;
            for(int j = 0; j < 8; ++j) a[r][j] -= f * a[c][j];
        }
    }
    mat4 r{};
    for(int i = 0; i < 4; ++i)
        for(int j = 0; j < 4; ++j) r[i][j] = a[i][4 + j];
    return r;
}

// glm::quatLookAt (right-handed): the rotation whose -Z axis is `direction`
quat quatLookAt(const vec3& direction, const vec3& up)
{
    const vec3 d = normalized(direction);
    const vec3 nd = {-d[0], -d[1], -d[2]};
    const vec3 right = normalized(cross(up, nd));
    const vec3 upv = cross(nd, right);
    std::array<std::array<double, 3>, 3> m{};
    for(int r = 0; r < 3; ++r) {
        m[r][0] = right[r];
        m[r][1] = upv[r];
        m[r][2] = nd[r];
    }
    return mat3ToQuat(m);
}

// transform.hpp:38-46: translate * toMat4(rotation) * scale
mat4 Transform::toMat4() const
{
    const auto R = quatToMat3(rotation);
    mat4 m = identity4();
    for(int r = 0; r < 3; ++r) {
        for(int c = 0; c < 3; ++c) m[r][c] = R[r][c] * scaling[c];
        m[r][3] = position[r];
    }
    return m;
}

bool Transform::isZeroVolume() const { return scaling[0] * scaling[1] * scaling[2] == 0.0; }

bool Transform::isIdentity() const
{
    return position == vec3{0.0, 0.0, 0.0} && rotation == quat{1.0, 0.0, 0.0, 0.0} && scaling == vec3{1.0, 1.0, 1.0};
}

Transform Transform::inverse() const
{
    Transform r;
    const double n = rotation[0] * rotation[0] + rotation[1] * rotation[1] + rotation[2] * rotation[2] + rotation[3] * rotation[3];
    for(int k = 0; k < 3; ++k) {
        r.position[k] = -position[k];
        r.scaling[k] = 1.0 / scaling[k];
    }
    r.rotation = {rotation[0] / n, -rotation[1] / n, -rotation[2] / n, -rotation[3] / n};
    return r;
}

quat quatFromEuler(const vec3& a)
{
    const double cx = std::cos(a[0] * 0.5), cy = std::cos(a[1] * 0.5), cz = std::cos(a[2] * 0.5);
    const double sx = std::sin(a[0] * 0.5), sy = std::sin(a[1] * 0.5), sz = std::sin(a[2] * 0.5);
    return {cx * cy * cz + sx * sy * sz, sx * cy * cz - cx * sy * sz, cx * sy * cz + sx * cy * sz, cx * cy * sz - sx * sy * cz};
}

void Transform::move(const vec3& t)
{
    for(int k = 0; k < 3; ++k) position[k] += t[k];
}

void Transform::rotate(double angle, const vec3& axis)
{
    vec3 a = axis;
    const double len = std::sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]);
    if(std::fabs(len - 1.0) > 0.001)
        for(double& c: a) c /= len;
    const double sn = std::sin(angle * 0.5);
    rotation = quatMul(rotation, {std::cos(angle * 0.5), a[0] * sn, a[1] * sn, a[2] * sn});
}

void Transform::rotate(const vec3& eulerAngles) { rotation = quatMul(quatFromEuler(eulerAngles), rotation); }

void Transform::rotateAround(const vec3& pivot, const vec3& eulerAngles)
{
    move({-pivot[0], -pivot[1], -pivot[2]});
    rotate(eulerAngles);
    move(pivot);
}

void Transform::lookAt(const vec3& target)
{
    rotation = quatLookAt(normalized({target[0] - position[0], target[1] - position[1], target[2] - position[2]}));
}

void Transform::scale(double f)
{
    for(double& c: scaling) c *= f;
}

void Transform::scale(const vec3& f)
{
    for(int k = 0; k < 3; ++k) scaling[k] *= f[k];
}

Transform interpolate(const Transform& x, const Transform& y, double factor)
{
    const double f = std::min(std::max(factor, 0.0), 1.0);
    Transform r;
    for(int k = 0; k < 3; ++k) {
        r.position[k] = x.position[k] * (1.0 - f) + y.position[k] * f;
        r.scaling[k] = x.scaling[k] * (1.0 - f) + y.scaling[k] * f;
    }
    for(int k = 0; k < 4; ++k) r.rotation[k] = x.rotation[k] * (1.0 - f) + y.rotation[k] * f;
    return r;
}

// transform.hpp:113-120
Transform Transform::operator*(const Transform& y) const
{
    const auto R = quatToMat3(rotation);
    Transform t;
    for(int r = 0; r < 3; ++r) {
        double s = 0.0;
        for(int c = 0; c < 3; ++c) s += R[r][c] * (scaling[c] * y.position[c]);
        t.position[r] = s + position[r];
    }
    t.rotation = quatMul(rotation, y.rotation);
    for(int k = 0; k < 3; ++k) t.scaling[k] = scaling[k] * y.scaling[k];
    return t;
}

// entity.cpp:180-270
void Entity::setTransform(const Transform& t)
{
    invalidateChildrenCachedParentTransform();
    m_transform = t;
}

void Entity::move(const vec3& translation)
{
    invalidateChildrenCachedParentTransform();
    m_transform.move(translation);
}

void Entity::moveTo(const vec3& position)
{
    invalidateChildrenCachedParentTransform();
    m_transform.position = position;
}

void Entity::rotate(double angle, const vec3& axis)
{
    invalidateChildrenCachedParentTransform();
    m_transform.rotate(angle, axis);
}

void Entity::rotate(const vec3& eulerAngles)
{
    invalidateChildrenCachedParentTransform();
    m_transform.rotate(eulerAngles);
}

void Entity::rotateAround(const vec3& pivot, const vec3& eulerAngles)
{
    invalidateChildrenCachedParentTransform();
    m_transform.rotateAround(pivot, eulerAngles);
}

void Entity::lookAt(const vec3& target)
{
    invalidateChildrenCachedParentTransform();
    m_transform.lookAt(target);
}

void Entity::scale(double factor)
{
    invalidateChildrenCachedParentTransform();
    m_transform.scale(factor);
}

void Entity::scale(const vec3& factors)
{
    invalidateChildrenCachedParentTransform();
    m_transform.scale(factors);
}

void Entity::setParent(const Entity* parent)
{
    invalidateCachedParentTransform();
    m_parent = parent;
}

void Entity::invalidateCachedParentTransform()
{
    m_cachedParentTransform.reset();
    invalidateChildrenCachedParentTransform();
}

void Entity::invalidateChildrenCachedParentTransform()
{
    for(const auto& c: children) c->invalidateCachedParentTransform();
}

// entity.cpp:187-199
Transform Entity::parentTransform() const
{
    if(!m_cachedParentTransform) m_cachedParentTransform = m_parent ? m_parent->globalTransform() : Transform();
    return *m_cachedParentTransform;
}

Transform Entity::globalTransform() const { return parentTransform() * m_transform; }

Entity& Entity::addChild(std::shared_ptr<Entity> e)
{
    e->setParent(this);
    children.push_back(std::move(e));
    return *children.back();
}

// entity.hpp:67-84
void Entity::forEachEntity(const std::function<bool(const Entity&)>& f) const
{
    if(!f(*this)) return;
    for(const auto& c: children) static_cast<const Entity&>(*c).forEachEntity(f);
}

void Entity::forEachEntity(const std::function<bool(Entity&)>& f)
{
    if(!f(*this)) return;
    for(const auto& c: children) c->forEachEntity(f);
}

// entity.cpp:300-332
bool EntityAnimation::update(double deltaTime, Entity& target)
{
    m_animationTime += deltaTime;
    return runAnimation(target);
}

bool ScaleAnimation::runAnimation(Entity& target)
{
    const double t = m_animationTime / m_duration;
    const double f = std::min(std::max(t, 0.0), 1.0);
    Transform tr = target.transform();
    for(int k = 0; k < 3; ++k) tr.scaling[k] = m_startScale[k] * (1.0 - f) + m_endScale[k] * f;
    target.setTransform(tr);
    return t <= 1.0;
}

void AnimatableEntity::update(double deltaTime)
{
    if(!m_animation) return;
    if(!m_animation->update(deltaTime, *this)) {
        m_animation.reset();
        if(m_animationFinisher) (*m_animationFinisher)();
        m_animationFinisher.reset();
    }
}

void Camera::lookAt(const vec3& eye, const vec3& target)
{
    Transform t = transform();
    t.position = eye;
    t.rotation = quatLookAt({target[0] - eye[0], target[1] - eye[1], target[2] - eye[2]});
    setTransform(t);
}

// camera.cpp:34-47: glm::perspectiveRH_ZO (depth zero to one) with the
// Y axis flipped for Vulkan's clip space.
mat4 Camera::projection(uint32_t width, uint32_t height) const
{
    double aspect = (double)((float)width / (float)height);
    if(!std::isfinite(aspect)) aspect = 16.0 / 9.0;
    const double f = (double)(float)(fovDeg * (3.14159265358979323846 / 180.0)); // glm::radians
    const double th = std::tan(f / 2.0);
    mat4 p{};
    p[0][0] = 1.0 / (aspect * th);
    p[1][1] = -(1.0 / th);
    p[2][2] = zFar / (zNear - zFar);
    p[3][2] = -1.0;
    p[2][3] = -(zFar * zNear) / (zFar - zNear);
    return p;
}

void Scene::update(double deltaTime)
{
    root->forEachEntity([deltaTime](Entity& e) {
        if(auto* a = dynamic_cast<AnimatableEntity*>(&e)) a->update(deltaTime);
        return true;
    });
}

std::vector<std::shared_ptr<render::Model>> Scene::models() const
{
    std::vector<std::shared_ptr<render::Model>> out;
    root->forEachEntity([&](const Entity& e) {
        if(e.model && std::find(out.begin(), out.end(), e.model) == out.end()) out.push_back(e.model);
        return true;
    });
    return out;
}

namespace render {

rgrt_material defaultMaterial()
{
    rgrt_material m{};
    m.diffuse[0] = 1.0f;
    m.diffuse[1] = 0.0f;
    m.diffuse[2] = 1.0f;
    m.transparency = 0.0f;
    m.specular[0] = m.specular[1] = m.specular[2] = 1.0f;
    m.reflectivity = 0.0f;
    m.roughness = 0.0f;
    m.ior = 1.0f;
    m.effectId = 0;
    m.rayConsumption = 1;
    m.emission = 0.0f;
    return m;
}

// acceleration_structure.cpp:140-194 (the range info of :167-170)
BottomLevelAS::BottomLevelAS(rgrt_ctx* ctx, const Mesh& mesh)
{
    rgrt_blas_desc d{};
    d.firstVertex = mesh.vertexBufferRef.offsetInElements();
    d.indexOffsetBytes = mesh.indexBufferRef.offsetInElements() * 4u;
    d.numFaces = mesh.numFaces();
    d.maxVertex = (uint32_t)mesh.vertices.size();
    ok(ctx, rgrt_build_blas(ctx, &d, &handle), "BottomLevelAS");
}

std::vector<rgrt_instance> TopLevelAS::collectInstances(const Scene& scene)
{
    std::vector<rgrt_instance> instances;
    scene.root->forEachEntity([&](const Entity& entity) {
        if(!entity.isVisible()) return false;          // do not descend
        if(entity.transform().isZeroVolume()) return false;
        if(!entity.model) return true;                  // children may still render
        if(!entity.model->mesh || !entity.model->mesh->bottomLevelAS)
            throw Error(RGRT_ERR_NOT_READY, "TopLevelAS: entity '" + entity.name + "' has no bottom-level structure");
        rgrt_instance in{};
        const mat4 m = entity.globalTransform().toMat4();
        for(int r = 0; r < 3; ++r)
            for(int c = 0; c < 4; ++c) in.transform[r * 4 + c] = (float)m[r][c];
        in.customIndex = (uint32_t)instances.size();
        in.mask = 0xffu;
        in.flags = RGRT_INSTANCE_TRIANGLE_CULL_DISABLE;
        in.blas = entity.model->mesh->bottomLevelAS->handle;
        in.offsets.vertexBufferOffset = entity.model->mesh->vertexBufferRef.offsetInElements();
        in.offsets.indexBufferOffset = entity.model->mesh->indexBufferRef.offsetInElements();
        in.offsets.materialBufferOffset = entity.model->materialBufferRef.offsetInElements();
        instances.push_back(in);
        return true;
    });
    return instances;
}

TopLevelAS::TopLevelAS(rgrt_ctx* ctx, const Scene& scene) : m_instances(collectInstances(scene))
{
    ok(ctx, rgrt_build_tlas(ctx, m_instances.data(), (uint32_t)m_instances.size()), "TopLevelAS");
}

bool TopLevelAS::sameLayout(const std::vector<rgrt_instance>& a, const std::vector<rgrt_instance>& b)
{
    if(a.size() != b.size()) return false;
    for(size_t i = 0; i < a.size(); ++i)
        if(a[i].blas != b[i].blas) return false;
    return true;
}

bool TopLevelAS::update(rgrt_ctx* ctx, const Scene& scene)
{
    std::vector<rgrt_instance> next = collectInstances(scene);
    const bool same = sameLayout(next, m_instances);
    m_instances = std::move(next);
    const uint32_t n = (uint32_t)m_instances.size();
    if(same) ok(ctx, rgrt_refit_tlas(ctx, m_instances.data(), n), "TopLevelAS::update");
    else ok(ctx, rgrt_build_tlas(ctx, m_instances.data(), n), "TopLevelAS::update");
    return same;
}

// raytracer.cpp:54-74: one structure per distinct mesh (the reference
// builds one per model; models sharing a mesh get identical trees)
void Raytracer::setupBottomLevelAS(const Scene& scene)
{
    for(const auto& model: scene.models())
        if(model->mesh && !model->mesh->bottomLevelAS) model->mesh->bottomLevelAS = std::make_shared<BottomLevelAS>(m_ctx, *model->mesh);
}

// raytracer.cpp:76-85: rebuilt every frame (refitted when refitTopLevelAS is
// set and only transforms changed)
void Raytracer::setupTopLevelAS(const Scene& scene)
{
    if(refitTopLevelAS && m_topLevelAS) m_topLevelAS->update(m_ctx, scene);
    else m_topLevelAS = std::make_unique<TopLevelAS>(m_ctx, scene);
}

void Raytracer::updateRenderTarget(const rgrt_ubo& uniformBuffer) { ok(m_ctx, rgrt_set_frame(m_ctx, &uniformBuffer), "updateRenderTarget"); }

// raytracer.cpp:87-147 + selectResultImage (:379-392)
rgrt_image Raytracer::doRaytracing()
{
    ok(m_ctx, rgrt_render(m_ctx, useFXAA ? 1 : 0), "doRaytracing");
    return RGRT_IMAGE_FINAL;
}

rgrt_image Raytracer::doFrame(const Scene& scene, const rgrt_ubo& uniformBuffer, bool graph)
{
    std::vector<rgrt_instance> inst = TopLevelAS::collectInstances(scene);
    const bool refit = refitTopLevelAS && m_topLevelAS && TopLevelAS::sameLayout(inst, m_topLevelAS->instances());
    const uint32_t flags =
        (useFXAA ? RGRT_FRAME_FXAA : 0u) | (refit ? RGRT_FRAME_REFIT_TLAS : 0u) | (graph ? RGRT_FRAME_GRAPH : 0u);
    ok(m_ctx, rgrt_frame(m_ctx, &uniformBuffer, inst.data(), (uint32_t)inst.size(), flags), "doFrame");
    m_topLevelAS = std::make_unique<TopLevelAS>(std::move(inst));
    return RGRT_IMAGE_FINAL;
}

void packModels(const std::vector<std::shared_ptr<Model>>& models, std::vector<rgrt_vertex>& vertices,
                std::vector<uint32_t>& indices, std::vector<rgrt_material>& materials)
{
    std::vector<Mesh*> meshes;
    for(const auto& m: models)
        if(m->mesh && std::find(meshes.begin(), meshes.end(), m->mesh.get()) == meshes.end()) meshes.push_back(m->mesh.get());
    vertices.clear();
    indices.clear();
    materials.clear();
    for(Mesh* mesh: meshes) {
        mesh->vertexBufferRef.offset = (uint32_t)vertices.size();
        mesh->indexBufferRef.offset = (uint32_t)indices.size();
        vertices.insert(vertices.end(), mesh->vertices.begin(), mesh->vertices.end());
        indices.insert(indices.end(), mesh->indices.begin(), mesh->indices.end());
    }
    for(const auto& m: models) {
        m->materialBufferRef.offset = (uint32_t)materials.size();
        materials.insert(materials.end(), m->materials.begin(), m->materials.end());
    }
}

rgrt_ubo cameraUbo(const Camera& camera, uint32_t width, uint32_t height, const vec3& lightDir, int32_t numSamples,
                   int32_t maxRecursions)
{
    rgrt_ubo u;
    std::memset(&u, 0, sizeof(u));
    const mat4 viewInverse = camera.globalTransform().toMat4();
    const mat4 projInverse = inverse(camera.projection(width, height));
    for(int c = 0; c < 4; ++c)
        for(int r = 0; r < 4; ++r) {
            u.viewInverse[c * 4 + r] = (float)viewInverse[r][c]; // column-major
            u.projInverse[c * 4 + r] = (float)projInverse[r][c];
        }
    u.clearColor[0] = u.clearColor[1] = u.clearColor[2] = 0.2f;
    u.numSamples = numSamples;
    const vec3 l = normalized(lightDir);
    for(int k = 0; k < 3; ++k) u.lightDir[k] = (float)l[k];
    u.maxRecursions = maxRecursions;
    u.time = 0.0f;
    u.showAlpha = 0;
    return u;
}

RenderSystem::RenderSystem(uint32_t width, uint32_t height, uint32_t flags, int32_t device)
    : m_width(width), m_height(height), m_flags(flags), m_device(device)
{
    rgrt_config cfg{device, width, height, flags};
    const rgrt_status s = rgrt_create(&m_ctx, &cfg);
    if(s != RGRT_OK) {
        std::string msg = "RenderSystem: rgrt_create failed (status " + std::to_string(s) + ")";
        if(m_ctx) {
            msg += std::string(": ") + rgrt_last_error(m_ctx);
            rgrt_destroy(m_ctx);
        }
        throw Error(s, msg);
    }
    m_raytracer = std::make_unique<Raytracer>(m_ctx);
    m_last = m_ctx;
    try {
        ok(m_ctx, rgrt_get_tuning(m_ctx, &m_defaultTuning), "RenderSystem");
        applyTuning();
    }
    catch(...) {
        m_raytracer.reset();
        rgrt_destroy(m_ctx);
        throw;
    }
}

void RenderSystem::applyTuning()
{
    rgrt_tuning t = m_defaultTuning;
    if(m_extra.empty()) {
        t.frameBands = 3;
        t.wavefrontPasses = 5;
        t.traceBlocksPerCU = 4;
        t.shadeBlocksPerCU = 8;
        t.drainBlocksPerCU = 2;
    }
    ok(m_ctx, rgrt_set_tuning(m_ctx, &t), "RenderSystem tuning");
    for(rgrt_ctx* c: m_extra) ok(c, rgrt_set_tuning(c, &t), "RenderSystem tuning");
}

RenderSystem::~RenderSystem()
{
    m_extraRt.clear();
    for(rgrt_ctx* c: m_extra) rgrt_destroy(c);
    m_raytracer.reset();
    if(m_ctx) rgrt_destroy(m_ctx);
}

void RenderSystem::setFramesInFlight(uint32_t n)
{
    if(n < 1) throw Error(RGRT_ERR_INVALID_ARG, "setFramesInFlight: n must be at least 1");
    m_extraRt.clear();
    for(rgrt_ctx* c: m_extra) rgrt_destroy(c);
    m_extra.clear();
    m_last = m_ctx;
    m_frame = 0;
    rgrt_config cfg{m_device, m_width, m_height, m_flags};
    for(uint32_t k = 1; k < n; ++k) {
        rgrt_ctx* c = nullptr;
        ok(m_ctx, rgrt_create_shared(m_ctx, &cfg, &c), "setFramesInFlight");
        m_extra.push_back(c);
        m_extraRt.push_back(std::make_unique<Raytracer>(c));
    }
    applyTuning();
}

void RenderSystem::setupModelBuffers(const Scene& scene)
{
    std::vector<rgrt_vertex> vertices;
    std::vector<uint32_t> indices;
    std::vector<rgrt_material> materials;
    const auto models = scene.models();
    for(const auto& m: models)
        if(m->mesh) m->mesh->bottomLevelAS.reset(); // buffers change: trees are rebuilt
    m_raytracer->resetTopLevelAS(); // new geometry drops every sharing context's TLAS
    for(auto& rt: m_extraRt) rt->resetTopLevelAS();
    packModels(models, vertices, indices, materials);
    ok(m_ctx, rgrt_upload_geometry(m_ctx, vertices.data(), vertices.size(), indices.data(), indices.size()), "setupModelBuffers");
    ok(m_ctx, rgrt_upload_materials(m_ctx, materials.data(), (uint32_t)materials.size()), "setupModelBuffers");
    m_raytracer->setupBottomLevelAS(scene);
}

rgrt_ubo RenderSystem::updateUniformBuffer(const Camera& camera, int32_t maxRecursions, int32_t numSamples, const vec3& lightDir)
{
    const rgrt_ubo u = cameraUbo(camera, m_width, m_height, lightDir, numSamples, maxRecursions);
    m_raytracer->updateRenderTarget(u);
    m_maxRecursions = maxRecursions;
    m_numSamples = numSamples;
    m_lightDir = lightDir;
    return u;
}

rgrt_image RenderSystem::renderFrame(const Scene& scene, bool graph)
{
    const rgrt_ubo u = cameraUbo(*scene.camera, m_width, m_height, m_lightDir, m_numSamples, m_maxRecursions);
    const size_t k = (size_t)(m_frame++ % framesInFlight());
    Raytracer& rt = k == 0 ? *m_raytracer : *m_extraRt[k - 1];
    rt.useFXAA = m_raytracer->useFXAA; // the first ray tracer's settings hold for all
    rt.refitTopLevelAS = m_raytracer->refitTopLevelAS;
    const rgrt_image img = rt.doFrame(scene, u, graph);
    m_last = rt.context();
    return img;
}

rgrt_image RenderSystem::render(const Scene& scene)
{
    m_raytracer->setupTopLevelAS(scene);
    m_last = m_ctx;
    return m_raytracer->doRaytracing();
}

std::vector<uint16_t> RenderSystem::readImage(rgrt_image which) const
{
    std::vector<uint16_t> out((size_t)m_width * m_height * 4);
    ok(m_last, rgrt_read_image(m_last, which, out.data(), out.size() * 2), "readImage");
    return out;
}

std::vector<uint8_t> RenderSystem::readImageRGBA8(rgrt_image which, rgrt_format format) const
{
    std::vector<uint8_t> out((size_t)m_width * m_height * 4);
    ok(m_last, rgrt_read_image_rgba8(m_last, which, format, out.data(), out.size()), "readImageRGBA8");
    return out;
}

void writePPM(const std::string& path, const std::vector<uint8_t>& rgba, uint32_t width, uint32_t height)
{
    if(rgba.size() != (size_t)width * height * 4) throw Error(RGRT_ERR_INVALID_ARG, "writePPM: size mismatch");
    std::FILE* f = std::fopen(path.c_str(), "wb");
    if(!f) throw Error(RGRT_ERR_INVALID_ARG, "writePPM: cannot open " + path);
    std::fprintf(f, "P6\n%u %u\n255\n", width, height);
    std::vector<uint8_t> row((size_t)width * 3);
    bool good = true;
    for(uint32_t y = 0; y < height && good; ++y) {
        for(uint32_t x = 0; x < width; ++x)
            for(int c = 0; c < 3; ++c) row[3 * x + c] = rgba[((size_t)y * width + x) * 4 + c];
        good = std::fwrite(row.data(), 1, row.size(), f) == row.size();
    }
    if(std::fclose(f) != 0 || !good) throw Error(RGRT_ERR_INVALID_ARG, "writePPM: write failed for " + path);
}

} // namespace render
} // namespace raygun
