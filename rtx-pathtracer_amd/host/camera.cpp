// Camera matrices as RayTracingApp::updateUniformBuffer builds them
// (src/RayTracingApp.cpp:559-585): view from CameraController::lookAt /
// getViewMatrix (src/CameraController.cpp:76-99), projection from
// getProjMatrix (src/CameraController.cpp:101-107: perspective with a
// Y flip, zero-to-one depth), and the two inverses.
#include <cmath>
#include <cstring>

#include "scene.h"

namespace rtxh {

namespace {
struct D3 {
    double x, y, z;
};
D3 sub(D3 a, D3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
D3 crs(D3 a, D3 b) { return {a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y}; }
double dt(D3 a, D3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
D3 nrm(D3 a) {
    double l = std::sqrt(dt(a, a));
    return {a.x / l, a.y / l, a.z / l};
}

// General 4x4 inverse by cofactors (column-major in and out).
bool invert(const double m[16], double out[16]) {
    double inv[16];
    inv[0] = m[5] * m[10] * m[15] - m[5] * m[11] * m[14] - m[9] * m[6] * m[15] + m[9] * m[7] * m[14] +
             m[13] * m[6] * m[11] - m[13] * m[7] * m[10];
    inv[4] = -m[4] * m[10] * m[15] + m[4] * m[11] * m[14] + m[8] * m[6] * m[15] - m[8] * m[7] * m[14] -
             m[12] * m[6] * m[11] + m[12] * m[7] * m[10];
    inv[8] = m[4] * m[9] * m[15] - m[4] * m[11] * m[13] - m[8] * m[5] * m[15] + m[8] * m[7] * m[13] +
             m[12] * m[5] * m[11] - m[12] * m[7] * m[9];
    inv[12] = -m[4] * m[9] * m[14] + m[4] * m[10] * m[13] + m[8] * m[5] * m[14] - m[8] * m[6] * m[13] -
              m[12] * m[5] * m[10] + m[12] * m[6] * m[9];
    inv[1] = -m[1] * m[10] * m[15] + m[1] * m[11] * m[14] + m[9] * m[2] * m[15] - m[9] * m[3] * m[14] -
             m[13] * m[2] * m[11] + m[13] * m[3] * m[10];
    inv[5] = m[0] * m[10] * m[15] - m[0] * m[11] * m[14] - m[8] * m[2] * m[15] + m[8] * m[3] * m[14] +
             m[12] * m[2] * m[11] - m[12] * m[3] * m[10];
    inv[9] = -m[0] * m[9] * m[15] + m[0] * m[11] * m[13] + m[8] * m[1] * m[15] - m[8] * m[3] * m[13] -
             m[12] * m[1] * m[11] + m[12] * m[3] * m[9];
    inv[13] = m[0] * m[9] * m[14] - m[0] * m[10] * m[13] - m[8] * m[1] * m[14] + m[8] * m[2] * m[13] +
              m[12] * m[1] * m[10] - m[12] * m[2] * m[9];
    inv[2] = m[1] * m[6] * m[15] - m[1] * m[7] * m[14] - m[5] * m[2] * m[15] + m[5] * m[3] * m[14] +
             m[13] * m[2] * m[7] - m[13] * m[3] * m[6];
    inv[6] = -m[0] * m[6] * m[15] + m[0] * m[7] * m[14] + m[4] * m[2] * m[15] - m[4] * m[3] * m[14] -
             m[12] * m[2] * m[7] + m[12] * m[3] * m[6];
    inv[10] = m[0] * m[5] * m[15] - m[0] * m[7] * m[13] - m[4] * m[1] * m[15] + m[4] * m[3] * m[13] +
              m[12] * m[1] * m[7] - m[12] * m[3] * m[5];
    inv[14] = -m[0] * m[5] * m[14] + m[0] * m[6] * m[13] + m[4] * m[1] * m[14] - m[4] * m[2] * m[13] -
              m[12] * m[1] * m[6] + m[12] * m[2] * m[5];
    inv[3] = -m[1] * m[6] * m[11] + m[1] * m[7] * m[10] + m[5] * m[2] * m[11] - m[5] * m[3] * m[10] -
             m[9] * m[2] * m[7] + m[9] * m[3] * m[6];
    inv[7] = m[0] * m[6] * m[11] - m[0] * m[7] * m[10] - m[4] * m[2] * m[11] + m[4] * m[3] * m[10] +
             m[8] * m[2] * m[7] - m[8] * m[3] * m[6];
    inv[11] = -m[0] * m[5] * m[11] + m[0] * m[7] * m[9] + m[4] * m[1] * m[11] - m[4] * m[3] * m[9] -
              m[8] * m[1] * m[7] + m[8] * m[3] * m[5];
    inv[15] = m[0] * m[5] * m[10] - m[0] * m[6] * m[9] - m[4] * m[1] * m[10] + m[4] * m[2] * m[9] +
              m[8] * m[1] * m[6] - m[8] * m[2] * m[5];
    double det = m[0] * inv[0] + m[1] * inv[4] + m[2] * inv[8] + m[3] * inv[12];
    if (det == 0.0 || !std::isfinite(det)) return false;
    for (int i = 0; i < 16; ++i) out[i] = inv[i] / det;
    return true;
}
} // namespace

bool cameraMatrices(const rtxh_view &v, uint32_t w, uint32_t h, rtx_camera *cam) {
    if (w == 0 || h == 0) return false;
    D3 pos{v.origin[0], v.origin[1], v.origin[2]};
    D3 tgt{v.target[0], v.target[1], v.target[2]};
    D3 up{v.up[0], v.up[1], v.up[2]};
    // quatLookAt(dir, up): columns (right, up', -dir); the view matrix is its
    // transpose followed by the translation by -pos.
    D3 dir = nrm(sub(tgt, pos));
    D3 back{-dir.x, -dir.y, -dir.z};
    D3 right = crs(up, back);
    double rl = dt(right, right);
    if (rl < 1e-5) rl = 1e-5;
    double inv = 1.0 / std::sqrt(rl);
    right = {right.x * inv, right.y * inv, right.z * inv};
    D3 upv = crs(back, right);
    double view[16] = {0};
    // rows of the rotation are right, upv, back
    view[0] = right.x; view[4] = right.y; view[8] = right.z;
    view[1] = upv.x;   view[5] = upv.y;   view[9] = upv.z;
    view[2] = back.x;  view[6] = back.y;  view[10] = back.z;
    view[12] = -dt(right, pos);
    view[13] = -dt(upv, pos);
    view[14] = -dt(back, pos);
    view[15] = 1.0;

    const double pi = 3.14159265358979323846;
    double fovy = v.vfov * pi / 180.0;
    double aspect = (double)w / (double)h;
    double tanHalf = std::tan(fovy / 2.0);
    const double zNear = 0.1, zFar = 1000.0;
    double proj[16] = {0};
    proj[0] = 1.0 / (aspect * tanHalf);
    proj[5] = -(1.0 / tanHalf); // proj[1][1] *= -1
    proj[10] = zFar / (zNear - zFar);
    proj[11] = -1.0;
    proj[14] = -(zFar * zNear) / (zFar - zNear);

    double viewInv[16], projInv[16];
    if (!invert(view, viewInv) || !invert(proj, projInv)) return false;
    for (int i = 0; i < 16; ++i) {
        cam->view[i] = (float)view[i];
        cam->proj[i] = (float)proj[i];
        cam->viewInverse[i] = (float)viewInv[i];
        cam->projInverse[i] = (float)projInv[i];
    }
    return true;
}

} // namespace rtxh
