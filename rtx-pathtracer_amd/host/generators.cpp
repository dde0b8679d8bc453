// Seeded procedural scenes for the benchmark workloads.  The reference
// loads scenes from Mitsuba XML / JSON files only; these generators build
// the same kinds of objects the loader produces (one model + instance per
// shape, area lights from emissive faces, analytic spheres) so everything
// downstream runs the reference's code paths.
#include <cmath>

#include "scene.h"

namespace rtxh {

// 5 walls (10 triangles) + 2 axis-aligned boxes (24) + a quad light (2),
// plus a dielectric sphere r=0.3, IOR 1.5.  Z is up.
bool generateCornell(uint64_t /*seed*/, Scene &s, std::string &err) {
    (void)err;
    int red = s.addMaterial(makeDiffuse({0.63f, 0.065f, 0.05f}));
    int green = s.addMaterial(makeDiffuse({0.14f, 0.45f, 0.09f}));
    int white = s.addMaterial(makeDiffuse({0.73f, 0.73f, 0.73f}));
    int light = s.addMaterial(makeLight({0.73f, 0.73f, 0.73f}, {17.0f, 12.0f, 4.0f}));
    int glass = s.addMaterial(makeDielectric(1.5f));

    Model walls; // box x in [-1,1], y in [-1,1] (open at y=-1), z in [0,2]; normals inward
    addQuad(walls, {-1, -1, 0}, {1, -1, 0}, {1, 1, 0}, {-1, 1, 0}, white);  // floor (+z)
    addQuad(walls, {-1, -1, 2}, {-1, 1, 2}, {1, 1, 2}, {1, -1, 2}, white);  // ceiling (-z)
    addQuad(walls, {-1, 1, 0}, {1, 1, 0}, {1, 1, 2}, {-1, 1, 2}, white);    // back (-y)
    addQuad(walls, {-1, -1, 0}, {-1, 1, 0}, {-1, 1, 2}, {-1, -1, 2}, red);  // left (+x)
    addQuad(walls, {1, -1, 0}, {1, -1, 2}, {1, 1, 2}, {1, 1, 0}, green);    // right (-x)
    s.addModelInstance(std::move(walls));

    Model boxes;
    addBox(boxes, {0.05f, -0.6f, 0.0f}, {0.65f, 0.0f, 0.6f}, white);   // short box
    addBox(boxes, {-0.7f, 0.0f, 0.0f}, {-0.1f, 0.6f, 1.2f}, white);    // tall box
    s.addModelInstance(std::move(boxes));

    Model lamp; // facing down, just below the ceiling
    addQuad(lamp, {-0.25f, -0.25f, 1.98f}, {-0.25f, 0.25f, 1.98f}, {0.25f, 0.25f, 1.98f},
            {0.25f, -0.25f, 1.98f}, light);
    s.addModelInstance(std::move(lamp));

    s.addSphere({0.35f, -0.3f, 0.9f}, 0.3f, glass); // resting on the short box

    s.view = rtxh_view{{0.0f, -3.8f, 1.0f}, {0.0f, 0.0f, 1.0f}, {0.0f, 0.0f, 1.0f}, 40.0f};
    s.buildLightTables();
    return true;
}

// A 20 x 10 x 20 m room (y up) holding exactly `triangles` triangles:
// 12 for the shell, 8 for four ceiling area lights, the rest random
// boxes (12 triangles) and 24-sided columns (48 side + 2*22 cap triangles)
// at uniform positions.  Materials per object: 50% Lambert (albedo
// U(0.1,0.8)), 30% GGX rough conductor (alpha U(0.05,0.5)), 20% dielectric (1.5).
bool generateRoom(uint64_t seed, uint32_t triangles, Scene &s, std::string &err) {
    if (triangles == 0) triangles = 262144;
    const uint32_t fixedTris = 12 + 8;
    if (triangles < fixedTris + 104 || (triangles - fixedTris) % 4 != 0) {
        err = "room scene: triangle count must be >= 124 and 20 + a multiple of 4";
        return false;
    }
    // 12*nb + 92*nc = T'  <=>  3*nb + 23*nc = T'/4
    uint32_t q = (triangles - fixedTris) / 4;
    uint32_t nc = q / 46; // about half the triangles in columns
    while ((q - 23 * nc) % 3 != 0) nc--;
    uint32_t nb = (q - 23 * nc) / 3;

    Pcg32 rng(seed, 0x5eedu);
    int shellMat = s.addMaterial(makeDiffuse({0.6f, 0.6f, 0.6f}));
    Model shell;
    const float X = 10.0f, H = 10.0f, Z = 10.0f;
    addQuad(shell, {-X, 0, -Z}, {-X, 0, Z}, {X, 0, Z}, {X, 0, -Z}, shellMat);   // floor (+y)
    addQuad(shell, {-X, H, -Z}, {X, H, -Z}, {X, H, Z}, {-X, H, Z}, shellMat);   // ceiling (-y)
    addQuad(shell, {-X, 0, -Z}, {X, 0, -Z}, {X, H, -Z}, {-X, H, -Z}, shellMat); // z=-Z wall
    addQuad(shell, {-X, 0, Z}, {-X, H, Z}, {X, H, Z}, {X, 0, Z}, shellMat);     // z=+Z wall
    addQuad(shell, {-X, 0, -Z}, {-X, H, -Z}, {-X, H, Z}, {-X, 0, Z}, shellMat); // x=-X wall
    addQuad(shell, {X, 0, -Z}, {X, 0, Z}, {X, H, Z}, {X, H, -Z}, shellMat);     // x=+X wall

    auto randomMaterial = [&]() {
        float u = rng.uniform();
        if (u < 0.5f) {
            float r = rng.uniform(0.1f, 0.8f), g = rng.uniform(0.1f, 0.8f), b = rng.uniform(0.1f, 0.8f);
            return s.addMaterial(makeDiffuse({r, g, b}));
        } else if (u < 0.8f) {
            float alpha = rng.uniform(0.05f, 0.5f);
            float eta = rng.uniform(0.2f, 1.5f);
            float k = rng.uniform(2.0f, 7.0f);
            return s.addMaterial(makeRoughConductor(alpha, eta, k, RTX_DISTRIBUTION_GGX));
        }
        return s.addMaterial(makeDielectric(1.5f));
    };

    for (uint32_t i = 0; i < nb; ++i) {
        int mat = randomMaterial();
        Vec3 half(rng.uniform(0.02f, 0.12f), rng.uniform(0.02f, 0.12f), rng.uniform(0.02f, 0.12f));
        Vec3 c(rng.uniform(-X + half.x, X - half.x), rng.uniform(half.y, H - half.y),
               rng.uniform(-Z + half.z, Z - half.z));
        addBox(shell, c - half, c + half, mat);
    }
    for (uint32_t i = 0; i < nc; ++i) {
        int mat = randomMaterial();
        float radius = rng.uniform(0.02f, 0.08f);
        float height = rng.uniform(0.3f, 1.5f);
        Vec3 base(rng.uniform(-X + radius, X - radius), rng.uniform(0.0f, H - height),
                  rng.uniform(-Z + radius, Z - radius));
        addColumn(shell, base, radius, height, 24, mat);
    }
    s.addModelInstance(std::move(shell));

    int lightMat = s.addMaterial(makeLight({0.5f, 0.5f, 0.5f}, {15.0f, 15.0f, 15.0f}));
    const float lx[4] = {-5, 5, -5, 5}, lz[4] = {-5, -5, 5, 5};
    for (int i = 0; i < 4; ++i) {
        Model lamp; // 2 x 2 m, facing down
        float y = H - 0.02f;
        addQuad(lamp, {lx[i] - 1, y, lz[i] - 1}, {lx[i] + 1, y, lz[i] - 1}, {lx[i] + 1, y, lz[i] + 1},
                {lx[i] - 1, y, lz[i] + 1}, lightMat);
        s.addModelInstance(std::move(lamp));
    }

    s.view = rtxh_view{{0.0f, 5.0f, -9.8f}, {0.0f, 4.5f, 0.0f}, {0.0f, 1.0f, 0.0f}, 55.0f};
    s.buildLightTables();
    return true;
}

// Plates-and-lights test after scenes/veachMIS/veachMIS.xml:11-117: four
// glossy GGX plates (alpha 0.005, 0.02, 0.05, 0.2) lit by four spherical lights
// (r = 0.03, 0.1, 0.3, 0.9) of equal power, over a diffuse floor.
bool generatePlates(uint64_t /*seed*/, Scene &s, std::string &err) {
    (void)err;
    const float radii[4] = {0.03f, 0.1f, 0.3f, 0.9f};
    const float cx[4] = {-3.75f, -1.25f, 1.25f, 3.75f};
    const float alphas[4] = {0.005f, 0.02f, 0.05f, 0.2f};
    // plate corner pairs (y, z) of the near and far edges, x spans [-4, 4]
    const float py0[4] = {-2.70651f, -3.28825f, -3.73096f, -3.99615f};
    const float pz0[4] = {0.25609f, 1.36972f, 2.70046f, 4.0667f};
    const float py1[4] = {-2.08375f, -2.83856f, -3.43378f, -3.82069f};
    const float pz1[4] = {-0.526323f, 0.476536f, 1.74564f, 3.08221f};

    int black = s.addMaterial(makeDiffuse({0, 0, 0}));
    for (int i = 0; i < 4; ++i) {
        float radiance = 100.0f * (0.1f * 0.1f) / (radii[i] * radii[i]); // equal power
        rtx_material lm = s.materials[black];
        lm.type = RTX_MAT_LIGHT;
        lm.lightColor[0] = lm.lightColor[1] = lm.lightColor[2] = radiance;
        int m = s.addMaterial(lm);
        s.addSphere({cx[i], 0.0f, 0.0f}, radii[i], m);
    }
    for (int i = 0; i < 4; ++i) {
        int m = s.addMaterial(makeRoughConductor(alphas[i], 0.370f, 2.820f, RTX_DISTRIBUTION_GGX));
        Model plate;
        addQuad(plate, {4, py0[i], pz0[i]}, {4, py1[i], pz1[i]}, {-4, py1[i], pz1[i]},
                {-4, py0[i], pz0[i]}, m);
        s.addModelInstance(std::move(plate));
    }
    int floorMat = s.addMaterial(makeDiffuse({0.4f, 0.04f, 0.04f}));
    Model floor;
    addQuad(floor, {-10, -4.14615f, -10}, {-10, -4.14615f, 10}, {10, -4.14615f, 10},
            {10, -4.14615f, -10}, floorMat);
    addQuad(floor, {-10, -10, -2}, {10, -10, -2}, {10, 10, -2}, {-10, 10, -2}, floorMat);
    s.addModelInstance(std::move(floor));

    s.view = rtxh_view{{0.0f, 2.0f, 15.0f}, {0.0f, -2.0f, 2.5f}, {0.0f, 1.0f, 0.0f}, 28.0f};
    s.buildLightTables();
    return true;
}

// A closed diffuse room (6 x 6 x 3, albedo 0.8) whose only light is a small
// square lamp (radiance 50) lying on top of a black box under the ceiling:
// the lamp lights a spot on the ceiling and the upper walls, and everything
// below the box sees it only through them, so BSDF sampling finds the bright
// spot rarely and a guide that learns its direction pays off.
bool generateSpot(uint64_t /*seed*/, Scene &s, std::string &err) {
    (void)err;
    int wall = s.addMaterial(makeDiffuse({0.8f, 0.8f, 0.8f}));
    int black = s.addMaterial(makeDiffuse({0.0f, 0.0f, 0.0f}));
    int lamp = s.addMaterial(makeLight({0.0f, 0.0f, 0.0f}, {50.0f, 50.0f, 50.0f}));
    Model light;
    addQuad(light, {-0.2f, -0.2f, 2.51f}, {0.2f, -0.2f, 2.51f}, {0.2f, 0.2f, 2.51f}, {-0.2f, 0.2f, 2.51f}, lamp);
    s.addModelInstance(std::move(light));
    Model shade;
    addBox(shade, {-0.21f, -0.21f, 2.3f}, {0.21f, 0.21f, 2.505f}, black);
    s.addModelInstance(std::move(shade));
    Model room;
    addQuad(room, {-3, -3, 0}, {3, -3, 0}, {3, 3, 0}, {-3, 3, 0}, wall);  // floor (+z)
    addQuad(room, {-3, -3, 3}, {-3, 3, 3}, {3, 3, 3}, {3, -3, 3}, wall);  // ceiling (-z)
    addQuad(room, {-3, -3, 0}, {-3, 3, 0}, {-3, 3, 3}, {-3, -3, 3}, wall);
    addQuad(room, {3, -3, 0}, {3, -3, 3}, {3, 3, 3}, {3, 3, 0}, wall);
    addQuad(room, {-3, 3, 0}, {3, 3, 0}, {3, 3, 3}, {-3, 3, 3}, wall);
    addQuad(room, {-3, -3, 0}, {-3, -3, 3}, {3, -3, 3}, {3, -3, 0}, wall);
    s.addModelInstance(std::move(room));
    s.view = rtxh_view{{0.0f, -2.8f, 1.5f}, {0.0f, 0.0f, 1.1f}, {0.0f, 0.0f, 1.0f}, 70.0f};
    s.buildLightTables();
    return true;
}

} // namespace rtxh
