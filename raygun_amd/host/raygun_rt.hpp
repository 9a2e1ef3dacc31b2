// raygun_rt.hpp — C++ host side of Raygun's ray-tracing path over the rgrt C ABI.
//
// Mirrors the classes a Raygun renderer calls, with the reference's names and
// call order, as far as the ray-tracing path consumes them:
//
//   Transform          raygun/transform.hpp:29-131 (T * R * S, operator*,
//                      the editing helpers, interpolate)
//   Entity             raygun/entity.hpp:32-131, entity.cpp:180-270 (cached
//                      parent transform, dropped when an ancestor moves)
//   EntityAnimation / ScaleAnimation / AnimatableEntity
//                      raygun/entity.hpp:133-175, entity.cpp:300-332
//   Camera             raygun/camera.hpp:29-48, camera.cpp:34-47
//   Scene              raygun/scene.hpp:36-52; update() = the animation
//                      step of the main loop (raygun.cpp:122-126)
//   render::Mesh       raygun/render/mesh.hpp:30-52 (+ BufferRef offsets)
//   render::Model      raygun/render/model.hpp:33-42
//   render::BottomLevelAS / TopLevelAS
//                      raygun/render/acceleration_structure.hpp:34-73
//   render::Raytracer  raygun/render/raytracer.hpp:36-46
//   render::RenderSystem (setupModelBuffers, updateUniformBuffer, render)
//                      raygun/render/render_system.cpp:88-330
//
// Vulkan objects become rgrt handles; the swapchain blit, ImGui and present
// that follow doRaytracing in the reference are out of scope. Errors: the
// reference aborts through RAYGUN_FATAL; here a failing rgrt call throws
// raygun::Error carrying the library's status and message.
#pragma once

#include <array>
#include <cstdint>
#include <functional>
#include <memory>
#include <optional>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/rgrt.h"

namespace raygun {

struct Error : std::runtime_error {
    rgrt_status status;
    Error(rgrt_status s, const std::string& what) : std::runtime_error(what), status(s) {}
};

using vec3 = std::array<double, 3>;
using quat = std::array<double, 4>; // w, x, y, z (glm order)
using mat4 = std::array<std::array<double, 4>, 4>; // m[row][col], column vectors

mat4 identity4();
mat4 mul(const mat4& a, const mat4& b);
mat4 inverse(const mat4& m);
quat quatLookAt(const vec3& direction, const vec3& up = {0.0, 1.0, 0.0});

quat quatFromEuler(const vec3& angles); // glm quat(vec3 eulerAngles)

// transform.hpp:29-120
struct Transform {
    vec3 position = {0.0, 0.0, 0.0};
    quat rotation = {1.0, 0.0, 0.0, 0.0};
    vec3 scaling = {1.0, 1.0, 1.0};

    mat4 toMat4() const; // translate * toMat4(rotation) * scale
    bool isZeroVolume() const;
    bool isIdentity() const;
    Transform inverse() const; // component-wise, as transform.hpp:55-62 defines it
    Transform operator*(const Transform& y) const;

    // transform.hpp:71-95
    void move(const vec3& translation);
    void rotate(double angle, const vec3& axis); // rotation * angleAxis (glm::rotate)
    void rotate(const vec3& eulerAngles);        // quat(euler) * rotation
    // move(-pivot), rotate, move(pivot): only the orientation changes, as in the reference
    void rotateAround(const vec3& pivot, const vec3& eulerAngles);
    void lookAt(const vec3& target);
    void scale(double factor);
    void scale(const vec3& factors);
};

// transform.hpp:122-131: component-wise lerp with the factor clamped to [0, 1]
Transform interpolate(const Transform& x, const Transform& y, double factor);

namespace render {

struct BottomLevelAS;
struct RenderSystem;

// A range of elements in one of the packed scene buffers (gpu::BufferRef).
struct BufferRef {
    uint32_t offset = 0; // in elements
    uint32_t offsetInElements() const { return offset; }
};

// mesh.hpp:30-52
struct Mesh {
    std::string name;
    std::vector<rgrt_vertex> vertices;
    std::vector<uint32_t> indices; // mesh-local
    BufferRef vertexBufferRef, indexBufferRef;
    std::shared_ptr<BottomLevelAS> bottomLevelAS; // one tree per distinct mesh

    uint32_t numFaces() const { return (uint32_t)(indices.size() / 3); }
};

// model.hpp:33-42
struct Model {
    std::shared_ptr<Mesh> mesh;
    std::vector<rgrt_material> materials;
    BufferRef materialBufferRef;
};

rgrt_material defaultMaterial(); // gpu_material.def:11-23

} // namespace render

// entity.hpp:32-131
struct Entity {
    std::string name;
    std::shared_ptr<render::Model> model;
    bool visible = true;
    std::vector<std::shared_ptr<Entity>> children;

    explicit Entity(std::string n = "") : name(std::move(n)) {}
    virtual ~Entity() = default;

    const Transform& transform() const { return m_transform; }
    bool isVisible() const { return visible; }
    const Entity* parent() const { return m_parent; }

    // entity.cpp:180-270: every change drops the cached parent transform of
    // the descendants
    void setTransform(const Transform& t);
    void move(const vec3& translation);
    void moveTo(const vec3& position);
    void rotate(double angle, const vec3& axis);
    void rotate(const vec3& eulerAngles);
    void rotateAround(const vec3& pivot, const vec3& eulerAngles);
    void lookAt(const vec3& target);
    void scale(double factor);
    void scale(const vec3& factors);

    Transform parentTransform() const; // cached
    Transform globalTransform() const; // parentTransform() * m_transform
    Entity& addChild(std::shared_ptr<Entity> e);
    // pre-order; the callback returns false to skip the subtree
    void forEachEntity(const std::function<bool(const Entity&)>& f) const;
    void forEachEntity(const std::function<bool(Entity&)>& f);

  private:
    void setParent(const Entity* parent);
    void invalidateCachedParentTransform();
    void invalidateChildrenCachedParentTransform();

    Transform m_transform;
    const Entity* m_parent = nullptr;
    mutable std::optional<Transform> m_cachedParentTransform;
};

// entity.hpp:133-153, entity.cpp:300-320
struct EntityAnimation {
    virtual ~EntityAnimation() = default;
    bool update(double deltaTime, Entity& target); // false once finished

  protected:
    virtual bool runAnimation(Entity& target) = 0;
    double m_animationTime = 0.0;
};

struct ScaleAnimation : EntityAnimation {
    ScaleAnimation(double duration, const vec3& startScale, const vec3& endScale)
        : m_duration(duration), m_startScale(startScale), m_endScale(endScale)
    {
    }

  protected:
    bool runAnimation(Entity& target) override;

  private:
    double m_duration;
    vec3 m_startScale, m_endScale;
};

// entity.hpp:155-175, entity.cpp:322-332
struct AnimatableEntity : Entity {
    explicit AnimatableEntity(std::string n = "") : Entity(std::move(n)) {}
    virtual void update(double deltaTime);

    template<typename T>
    void setAnimation(const T& animation)
    {
        m_animation = std::make_unique<T>(animation);
    }
    void setAnimationFinisher(const std::function<void()>& f) { m_animationFinisher = f; }
    bool animating() const { return m_animation != nullptr; }

  protected:
    std::unique_ptr<EntityAnimation> m_animation;
    std::optional<std::function<void()>> m_animationFinisher;
};

// camera.hpp:29-48; the field of view is a parameter (reference: 45 degrees)
struct Camera : Entity {
    double fovDeg = 45.0, zNear = 0.1, zFar = 100.0;
    Camera() : Entity("camera") {}
    void lookAt(const vec3& eye, const vec3& target);
    mat4 projection(uint32_t width, uint32_t height) const; // perspectiveRH_ZO, [1][1] *= -1
};

// scene.hpp:36-52
struct Scene {
    std::shared_ptr<Entity> root = std::make_shared<Entity>("root");
    std::shared_ptr<Camera> camera = std::make_shared<Camera>();
    // distinct models in entity pre-order (first use)
    std::vector<std::shared_ptr<render::Model>> models() const;
    // the animation step of the main loop (raygun.cpp:122-126)
    void update(double deltaTime);
};

namespace render {

// acceleration_structure.hpp:55-73: one tree over a mesh's triangles.
struct BottomLevelAS {
    BottomLevelAS(rgrt_ctx* ctx, const Mesh& mesh);
    explicit BottomLevelAS(rgrt_blas h) : handle(h) {} // a tree built elsewhere
    rgrt_blas handle = 0;
};

// acceleration_structure.hpp:34-53 + instanceOffsetTable(): the DFS over
// the entity tree (acceleration_structure.cpp:55-85). Invisible and
// zero-volume subtrees are skipped, entities without a model are descended
// into; instance i gets customIndex i, mask 0xff, the transposed global
// transform's first three rows and its offset-table entry.
struct TopLevelAS {
    TopLevelAS(rgrt_ctx* ctx, const Scene& scene);
    // a TLAS the context built from these instances in another call (rgrt_frame)
    explicit TopLevelAS(std::vector<rgrt_instance> built) : m_instances(std::move(built)) {}
    // same instance count and BLAS per position: a refit suffices
    static bool sameLayout(const std::vector<rgrt_instance>& a, const std::vector<rgrt_instance>& b);
    static std::vector<rgrt_instance> collectInstances(const Scene& scene);
    const std::vector<rgrt_instance>& instances() const { return m_instances; }
    // Per-frame update without a rebuild (no counterpart in the reference):
    // when the scene yields the same instance count and BLASes (only
    // transforms changed) the tree is refitted (rgrt_refit_tlas); otherwise
    // it is rebuilt. Returns true for a refit.
    bool update(rgrt_ctx* ctx, const Scene& scene);

  private:
    std::vector<rgrt_instance> m_instances;
};

// raytracer.hpp:36-46
struct Raytracer {
    explicit Raytracer(rgrt_ctx* ctx) : m_ctx(ctx) {}

    void setupBottomLevelAS(const Scene& scene);
    void setupTopLevelAS(const Scene& scene);
    void updateRenderTarget(const rgrt_ubo& uniformBuffer);
    // trace + rough prepare + 10 x (H, V) blur + postprocess + FXAA + swap;
    // returns the image the reference's selectResultImage() would return
    rgrt_image doRaytracing();
    // setupTopLevelAS + doRaytracing with the frame constants as one
    // rgrt_frame call, launched as one hipGraph when graph is set
    rgrt_image doFrame(const Scene& scene, const rgrt_ubo& uniformBuffer, bool graph);

    bool useFXAA = true; // m_useFXAA (raytracer.hpp:80)
    // The reference rebuilds the TLAS every frame; with refitTopLevelAS set,
    // frames that only move instances refit it (TopLevelAS::update).
    bool refitTopLevelAS = false;
    // new geometry drops the context's TLAS: the next frame builds one
    void resetTopLevelAS() { m_topLevelAS.reset(); }
    rgrt_ctx* context() const { return m_ctx; }

  private:
    rgrt_ctx* m_ctx;
    std::unique_ptr<TopLevelAS> m_topLevelAS;
};

// render_system.cpp: the parts around the ray tracer.
struct RenderSystem {
    RenderSystem(uint32_t width, uint32_t height, uint32_t flags = 0, int32_t device = 0);
    ~RenderSystem();
    RenderSystem(const RenderSystem&) = delete;
    RenderSystem& operator=(const RenderSystem&) = delete;

    // packs the distinct meshes and the models' materials into the shared
    // buffers (render_system.cpp:192-223, 270-330) and builds the BLASes
    // (Raygun::finalizeLoadScene, raygun.cpp:276-295)
    void setupModelBuffers(const Scene& scene);
    // resetUniformBuffer + updateUniformBuffer (render_system.cpp:235-268)
    rgrt_ubo updateUniformBuffer(const Camera& camera, int32_t maxRecursions = 5, int32_t numSamples = 1,
                                 const vec3& lightDir = {0.4, -0.6, -0.8});
    // per frame: TLAS + doRaytracing (render_system.cpp:88-100)
    rgrt_image render(const Scene& scene);
    // the whole frame of render_system.cpp:88-100 (frame constants from the
    // scene camera with the settings of the last updateUniformBuffer, TLAS,
    // trace, post) as one rgrt_frame call; graph: one hipGraph launch
    rgrt_image renderFrame(const Scene& scene, bool graph = true);
    // Frames in flight (no counterpart in the reference, which renders one
    // frame at a time): n - 1 more contexts sharing the scene
    // (rgrt_create_shared); renderFrame then rotates over the n contexts, so
    // a frame starts while the previous ones still run. readImage and
    // readImageRGBA8 read the most recent renderFrame's images.
    // Launch shape (no counterpart in the reference): with one frame in
    // flight every frame is traced in three strips side by side
    // (rgrt_tuning.frameBands 3, 5 wavefront passes, 4/8/2 trace/shade/drain
    // blocks per CU, measured on the 1080p 2M-triangle scene); with more,
    // the contexts already run side by side and keep the library's one-band
    // default. Both constructor and setFramesInFlight set it on every context.
    void setFramesInFlight(uint32_t n);
    uint32_t framesInFlight() const { return 1u + (uint32_t)m_extra.size(); }
    std::vector<uint16_t> readImage(rgrt_image which) const; // rgba16f bits
    // the swapchain blit's conversion (render_system.cpp:102-147) + readback:
    // 4 bytes per pixel in the given 8-bit format
    std::vector<uint8_t> readImageRGBA8(rgrt_image which, rgrt_format format = RGRT_FORMAT_RGBA8_UNORM) const;

    Raytracer& raytracer() { return *m_raytracer; }
    rgrt_ctx* context() const { return m_ctx; }
    uint32_t width() const { return m_width; }
    uint32_t height() const { return m_height; }

  private:
    rgrt_ctx* m_ctx = nullptr;
    uint32_t m_width, m_height, m_flags;
    int32_t m_device;
    std::unique_ptr<Raytracer> m_raytracer;
    std::vector<rgrt_ctx*> m_extra;                    // frames in flight beyond the first
    std::vector<std::unique_ptr<Raytracer>> m_extraRt; // one per extra context
    uint64_t m_frame = 0;
    rgrt_ctx* m_last = nullptr; // context of the most recent renderFrame / render
    int32_t m_maxRecursions = 5, m_numSamples = 1;
    vec3 m_lightDir = {0.4, -0.6, -0.8};
    rgrt_tuning m_defaultTuning{}; // the library's launch shape, as rgrt_create left it
    void applyTuning();
};

// packing without a context: fills the buffer refs, returns the buffers
void packModels(const std::vector<std::shared_ptr<Model>>& models, std::vector<rgrt_vertex>& vertices,
                std::vector<uint32_t>& indices, std::vector<rgrt_material>& materials);

// Binary PPM (P6) of 4-byte RGBA pixels (alpha dropped).
void writePPM(const std::string& path, const std::vector<uint8_t>& rgba, uint32_t width, uint32_t height);

// UBO of a frame (render_system.cpp:235-268; matrices column-major)
rgrt_ubo cameraUbo(const Camera& camera, uint32_t width, uint32_t height, const vec3& lightDir, int32_t numSamples,
                   int32_t maxRecursions);

} // namespace render
} // namespace raygun
