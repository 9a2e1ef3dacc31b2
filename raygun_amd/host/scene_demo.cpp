// scene_demo.cpp — drives the C++ host mirror (raygun_rt.hpp) on a small
// scene with nested, hidden and zero-volume entities, shared meshes and three
// materials. The same scene is built in Python by tests/host_scene.py.
//
//   scene_demo collect                 print the TLAS instances, the packed
//                                      buffer offsets and the frame constants
//                                      as JSON (no GPU needed)
//   scene_demo render W H OUT          render one frame; write the final image
//                                      (rgba16f) to OUT and the frame constants
//                                      to OUT.ubo
//   scene_demo frames W H N OUT.ppm [IN_FLIGHT]
//                                      N animated frames through rgrt_frame
//                                      (one hipGraph per frame, TLAS refit);
//                                      the last one read back as 8-bit RGBA
//                                      and written as PPM, the same frames
//                                      traced eagerly compared byte for byte;
//                                      IN_FLIGHT contexts sharing the scene
//                                      take the frames in turn
//   scene_demo animate FRAMES          run FRAMES frames of the per-frame update
//                                      (a scale animation on cube2, the group
//                                      turning, cube4 moving) and print the
//                                      TLAS instances of every frame as JSON
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <string>

#include "raygun_rt.hpp"

using namespace raygun;
using render::Mesh;
using render::Model;

namespace {

rgrt_vertex vtx(float x, float y, float z, float nx, float ny, float nz)
{
    rgrt_vertex v{};
    v.position[0] = x;
    v.position[1] = y;
    v.position[2] = z;
    v.matIndex = 0;
    v.normal[0] = nx;
    v.normal[1] = ny;
    v.normal[2] = nz;
    return v;
}

// unit cube centred at the origin, 4 vertices and 2 triangles per face
std::shared_ptr<Mesh> cube()
{
    auto m = std::make_shared<Mesh>();
    m->name = "cube";
    for(int axis = 0; axis < 3; ++axis)
        for(int sgn = 0; sgn < 2; ++sgn) {
            const float s = sgn ? 1.0f : -1.0f;
            const int u = (axis + 1) % 3, w = (axis + 2) % 3;
            const uint32_t base = (uint32_t)m->vertices.size();
            for(int k = 0; k < 4; ++k) {
                float p[3], n[3] = {0.0f, 0.0f, 0.0f};
                p[axis] = 0.5f * s;
                p[u] = (k == 1 || k == 2) ? 0.5f : -0.5f;
                p[w] = (k >= 2) ? 0.5f : -0.5f;
                n[axis] = s;
                m->vertices.push_back(vtx(p[0], p[1], p[2], n[0], n[1], n[2]));
            }
            const uint32_t a[6] = {0, 1, 2, 0, 2, 3};
            for(uint32_t i: a) m->indices.push_back(base + i);
        }
    return m;
}

std::shared_ptr<Mesh> ground()
{
    auto m = std::make_shared<Mesh>();
    m->name = "ground";
    m->vertices = {vtx(-1, 0, -1, 0, 1, 0), vtx(1, 0, -1, 0, 1, 0), vtx(1, 0, 1, 0, 1, 0), vtx(-1, 0, 1, 0, 1, 0)};
    m->indices = {0, 2, 1, 0, 3, 2};
    return m;
}

quat axisAngle(double ax, double ay, double az, double angleDeg)
{
    const double l = std::sqrt(ax * ax + ay * ay + az * az);
    const double h = angleDeg * (3.14159265358979323846 / 180.0) / 2.0;
    const double s = std::sin(h);
    return {std::cos(h), ax / l * s, ay / l * s, az / l * s};
}

template<typename E = Entity>
std::shared_ptr<E> entity(const std::string& name, std::shared_ptr<Model> model, vec3 pos, quat rot = {1, 0, 0, 0},
                          vec3 scale = {1, 1, 1})
{
    auto e = std::make_shared<E>(name);
    e->model = std::move(model);
    Transform t;
    t.position = pos;
    t.rotation = rot;
    t.scaling = scale;
    e->setTransform(t);
    return e;
}

Scene demoScene(bool animated = false)
{
    auto cubeMesh = cube();
    auto groundMesh = ground();
    auto red = std::make_shared<Model>();
    red->mesh = cubeMesh;
    rgrt_material m = render::defaultMaterial();
    m.diffuse[0] = 0.8f;
    m.diffuse[1] = 0.1f;
    m.diffuse[2] = 0.1f;
    m.reflectivity = 0.5f;
    m.roughness = 0.3f;
    red->materials = {m};
    auto glass = std::make_shared<Model>();
    glass->mesh = cubeMesh;
    m = render::defaultMaterial();
    m.diffuse[0] = 0.9f;
    m.diffuse[1] = 0.9f;
    m.diffuse[2] = 1.0f;
    m.transparency = 0.9f;
    m.ior = 1.5f;
    glass->materials = {m};
    auto floor = std::make_shared<Model>();
    floor->mesh = groundMesh;
    m = render::defaultMaterial();
    m.diffuse[0] = m.diffuse[1] = m.diffuse[2] = 0.6f;
    m.reflectivity = 0.3f;
    m.roughness = 0.2f;
    m.effectId = 1;
    floor->materials = {m};

    Scene s;
    s.root->addChild(entity("ground", floor, {0, 0, 0}, {1, 0, 0, 0}, {20, 1, 20}));
    auto& group = s.root->addChild(entity("group", nullptr, {0, 0.5, 0}, axisAngle(0, 1, 0, 30)));
    group.addChild(entity("cube1", red, {-1.2, 0, 0}));
    if(animated) {
        auto c2 = entity<AnimatableEntity>("cube2", glass, {1.2, 0.2, 0}, {1, 0, 0, 0}, {0.8, 0.8, 0.8});
        c2->setAnimation(ScaleAnimation(1.0, {0.8, 0.8, 0.8}, {1.3, 1.3, 1.3}));
        group.addChild(c2);
    }
    else {
        group.addChild(entity("cube2", glass, {1.2, 0.2, 0}, {1, 0, 0, 0}, {0.8, 0.8, 0.8}));
    }
    auto& hidden = group.addChild(entity("hidden", red, {0, 3, 0}));
    hidden.visible = false;
    hidden.addChild(entity("under-hidden", glass, {0, 1, 0}));
    group.addChild(entity("flat", red, {0, 0, 2}, {1, 0, 0, 0}, {1, 0, 1}));
    s.root->addChild(entity("cube4", red, {0, 1.8, -1}, axisAngle(1, 1, 0, 40), {0.6, 0.6, 0.6}));
    s.camera->fovDeg = 60.0;
    s.camera->lookAt({0, 2, 6}, {0, 0.5, 0});
    return s;
}

void printFloats(const float* f, int n)
{
    std::printf("[");
    for(int i = 0; i < n; ++i) std::printf("%s%.9g", i ? ", " : "", (double)f[i]);
    std::printf("]");
}

Entity* findEntity(Entity& root, const std::string& name)
{
    Entity* found = nullptr;
    root.forEachEntity([&](Entity& e) {
        if(e.name == name) found = &e;
        return found == nullptr;
    });
    return found;
}

} // namespace

int main(int argc, char** argv)
{
    const std::string mode = argc > 1 ? argv[1] : "";
    try {
        Scene s = demoScene();
        if(mode == "collect") {
            std::vector<rgrt_vertex> v;
            std::vector<uint32_t> i;
            std::vector<rgrt_material> mats;
            const auto models = s.models();
            render::packModels(models, v, i, mats);
            // handles as a context would hand them out: distinct meshes in first use order
            rgrt_blas next = 0;
            for(const auto& mo: models)
                if(!mo->mesh->bottomLevelAS) mo->mesh->bottomLevelAS = std::make_shared<render::BottomLevelAS>(next++);
            const auto inst = render::TopLevelAS::collectInstances(s);
            std::printf("{\"vertices\": %zu, \"indices\": %zu, \"materials\": %zu,\n \"instances\": [", v.size(), i.size(),
                        mats.size());
            for(size_t k = 0; k < inst.size(); ++k) {
                const rgrt_instance& r = inst[k];
                std::printf("%s\n  {\"transform\": ", k ? "," : "");
                printFloats(r.transform, 12);
                std::printf(", \"customIndex\": %u, \"mask\": %u, \"flags\": %u, \"blas\": %u, \"offsets\": [%u, %u, %u]}",
                            r.customIndex, r.mask, r.flags, r.blas, r.offsets.vertexBufferOffset, r.offsets.indexBufferOffset,
                            r.offsets.materialBufferOffset);
            }
            const rgrt_ubo u = render::cameraUbo(*s.camera, 320, 180, {0.4, -0.6, -0.8}, 1, 3);
            std::printf("],\n \"viewInverse\": ");
            printFloats(u.viewInverse, 16);
            std::printf(",\n \"projInverse\": ");
            printFloats(u.projInverse, 16);
            std::printf(",\n \"lightDir\": ");
            printFloats(u.lightDir, 3);
            std::printf("}\n");
            return 0;
        }
        if(mode == "animate" && argc == 3) {
            Scene a = demoScene(true);
            const int frames = std::atoi(argv[2]);
            rgrt_blas next = 0;
            for(const auto& mo: a.models())
                if(!mo->mesh->bottomLevelAS) mo->mesh->bottomLevelAS = std::make_shared<render::BottomLevelAS>(next++);
            Entity* group = findEntity(*a.root, "group");
            Entity* cube4 = findEntity(*a.root, "cube4");
            int finished = 0;
            static_cast<AnimatableEntity*>(findEntity(*a.root, "cube2"))->setAnimationFinisher([&]() { ++finished; });
            std::printf("{\"frames\": [");
            for(int f = 0; f < frames; ++f) {
                a.update(0.25);
                group->rotate(vec3{0.0, 0.1, 0.0});
                cube4->move({0.05, 0.0, 0.0});
                const auto inst = render::TopLevelAS::collectInstances(a);
                std::printf("%s\n  [", f ? "," : "");
                for(size_t k = 0; k < inst.size(); ++k) {
                    std::printf("%s{\"blas\": %u, \"transform\": ", k ? ", " : "", inst[k].blas);
                    printFloats(inst[k].transform, 12);
                    std::printf("}");
                }
                std::printf("]");
            }
            std::printf("],\n \"finished\": %d}\n", finished);
            return 0;
        }
        if(mode == "frames" && (argc == 6 || argc == 7)) {
            const uint32_t W = (uint32_t)std::atoi(argv[2]), H = (uint32_t)std::atoi(argv[3]);
            const int frames = std::atoi(argv[4]);
            const uint32_t inFlight = argc == 7 ? (uint32_t)std::atoi(argv[6]) : 1u;
            Scene g = demoScene(true), e = demoScene(true);
            render::RenderSystem rg(W, H), re(W, H);
            rg.setFramesInFlight(inFlight);
            rg.setupModelBuffers(g);
            re.setupModelBuffers(e);
            rg.raytracer().refitTopLevelAS = true;
            rg.updateUniformBuffer(*g.camera, 3, 1); // the frame settings renderFrame uses
            std::vector<uint8_t> last;
            for(int f = 0; f < frames; ++f) {
                for(Scene* sc: {&g, &e}) {
                    sc->update(0.25);
                    findEntity(*sc->root, "group")->rotate(vec3{0.0, 0.1, 0.0});
                    findEntity(*sc->root, "cube4")->move({0.05, 0.0, 0.0});
                }
                rg.renderFrame(g, /*graph=*/true);
                re.updateUniformBuffer(*e.camera, 3, 1);
                re.render(e);
                last = rg.readImageRGBA8(RGRT_IMAGE_FINAL);
                if(last != re.readImageRGBA8(RGRT_IMAGE_FINAL)) {
                    std::fprintf(stderr, "frame %d: graph and eager frames differ\n", f);
                    return 1;
                }
            }
            render::writePPM(argv[5], last, W, H);
            std::printf("%d frames %ux%u, %u in flight, graph == eager\n", frames, W, H, inFlight);
            return 0;
        }
        if(mode == "render" && argc == 5) {
            const uint32_t W = (uint32_t)std::atoi(argv[2]), H = (uint32_t)std::atoi(argv[3]);
            render::RenderSystem rs(W, H);
            rs.setupModelBuffers(s);
            const rgrt_ubo ubo = rs.updateUniformBuffer(*s.camera, 3, 1);
            const rgrt_image img = rs.render(s);
            const std::vector<uint16_t> px = rs.readImage(img);
            std::ofstream(argv[4], std::ios::binary).write((const char*)px.data(), (std::streamsize)(px.size() * 2));
            std::ofstream(std::string(argv[4]) + ".ubo", std::ios::binary).write((const char*)&ubo, sizeof(ubo));
            std::printf("rendered %ux%u, %zu instances\n", W, H, render::TopLevelAS::collectInstances(s).size());
            return 0;
        }
        std::fprintf(stderr, "usage: scene_demo collect | render W H OUT | animate FRAMES | frames W H N OUT.ppm [IN_FLIGHT]\n");
        return 2;
    }
    catch(const Error& e) {
        std::fprintf(stderr, "scene_demo: %s\n", e.what());
        return 1;
    }
}
