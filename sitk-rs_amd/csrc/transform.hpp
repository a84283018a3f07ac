// Host-side parametric transforms of the registration driver: Translation,
// Euler3D and Affine (crates/sitk/src/transform/parametric.rs:357-631,
// 935-1180). Each reports its point map as one stage (matrix, offset) -- the
// fields its own transform_point evaluates -- and its parameter Jacobian.
#pragma once

#include <cmath>
#include <cstring>

#include "../../include/sitk_hip.h"

namespace sitkhip {

struct Transform {
    int kind = SITKHIP_TRANSFORM_EULER3D;
    int nparams = 6;
    double params[12] = {0};
    double center[3] = {0, 0, 0};
    bool compute_zyx = false;
    double matrix[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
    double offset[3] = {0, 0, 0};

    static void matmul3(const double a[9], const double b[9], double out[9]) {
        // crates/sitk/src/core/matrix.rs:24-40 (a zero a[r][k] is skipped)
        for (int i = 0; i < 9; ++i) out[i] = 0.0;
        for (int r = 0; r < 3; ++r)
            for (int k = 0; k < 3; ++k) {
                const double ark = a[r * 3 + k];
                if (ark == 0.0) continue;
                for (int c = 0; c < 3; ++c) out[r * 3 + c] += ark * b[k * 3 + c];
            }
    }
    static void mat_vec3(const double m[9], const double v[3], double out[3]) {
        for (int r = 0; r < 3; ++r) {
            double acc = 0.0;
            for (int c = 0; c < 3; ++c) acc += m[r * 3 + c] * v[c];
            out[r] = acc;
        }
    }

    // Rebuild matrix and offset from the parameters (Euler3D recompute,
    // parametric.rs:1030-1053; Affine compute_offset :513-519).
    void recompute() {
        if (kind == SITKHIP_TRANSFORM_TRANSLATION) {
            nparams = 3;
            const double I[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
            std::memcpy(matrix, I, sizeof(I));
            for (int d = 0; d < 3; ++d) offset[d] = params[d];
        } else if (kind == SITKHIP_TRANSFORM_EULER3D) {
            nparams = 6;
            const double cx = std::cos(params[0]), sx = std::sin(params[0]);
            const double cy = std::cos(params[1]), sy = std::sin(params[1]);
            const double cz = std::cos(params[2]), sz = std::sin(params[2]);
            const double rx[9] = {1.0, 0.0, 0.0, 0.0, cx, -sx, 0.0, sx, cx};
            const double ry[9] = {cy, 0.0, sy, 0.0, 1.0, 0.0, -sy, 0.0, cy};
            const double rz[9] = {cz, -sz, 0.0, sz, cz, 0.0, 0.0, 0.0, 1.0};
            double t[9];
            if (compute_zyx) { matmul3(rz, ry, t); matmul3(t, rx, matrix); }
            else { matmul3(rz, rx, t); matmul3(t, ry, matrix); }
            double mc[3];
            mat_vec3(matrix, center, mc);
            for (int i = 0; i < 3; ++i) offset[i] = params[3 + i] + center[i] - mc[i];
        } else {
            nparams = 12;
            std::memcpy(matrix, params, 9 * sizeof(double));
            double ac[3];
            mat_vec3(matrix, center, ac);
            for (int d = 0; d < 3; ++d) offset[d] = center[d] + params[9 + d] - ac[d];
        }
    }

    void set_parameters(const double *p) {
        std::memcpy(params, p, sizeof(double) * static_cast<size_t>(nparams));
        recompute();
    }

    void stage(double out[12]) const {
        std::memcpy(out, matrix, 9 * sizeof(double));
        std::memcpy(out + 9, offset, 3 * sizeof(double));
    }

    void apply(const double x[3], double out[3]) const {
        if (kind == SITKHIP_TRANSFORM_TRANSLATION) {
            for (int d = 0; d < 3; ++d) out[d] = x[d] + params[d];
            return;
        }
        double mx[3];
        mat_vec3(matrix, x, mx);
        for (int d = 0; d < 3; ++d) out[d] = mx[d] + offset[d];
    }

    // Row-major 3 x nparams Jacobian at point p (parametric.rs:453-461,
    // 594-607, 1134-1179).
    void jacobian(const double p[3], double *j) const {
        const int np = nparams;
        for (int i = 0; i < 3 * np; ++i) j[i] = 0.0;
        if (kind == SITKHIP_TRANSFORM_TRANSLATION) {
            for (int i = 0; i < 3; ++i) j[i * 3 + i] = 1.0;
        } else if (kind == SITKHIP_TRANSFORM_AFFINE) {
            for (int i = 0; i < 3; ++i) {
                for (int c = 0; c < 3; ++c) j[i * np + i * 3 + c] = p[c] - center[c];
                j[i * np + 9 + i] = 1.0;
            }
        } else {
            const double cx = std::cos(params[0]), sx = std::sin(params[0]);
            const double cy = std::cos(params[1]), sy = std::sin(params[1]);
            const double cz = std::cos(params[2]), sz = std::sin(params[2]);
            const double px = p[0] - center[0], py = p[1] - center[1], pz = p[2] - center[2];
            if (compute_zyx) {
                j[0] = (cz * sy * cx + sz * sx) * py + (-cz * sy * sx + sz * cx) * pz;
                j[6] = (sz * sy * cx - cz * sx) * py + (-sz * sy * sx - cz * cx) * pz;
                j[12] = (cy * cx) * py + (-cy * sx) * pz;
                j[1] = (-cz * sy) * px + (cz * cy * sx) * py + (cz * cy * cx) * pz;
                j[7] = (-sz * sy) * px + (sz * cy * sx) * py + (sz * cy * cx) * pz;
                j[13] = (-cy) * px + (-sy * sx) * py + (-sy * cx) * pz;
                j[2] = (-sz * cy) * px + (-sz * sy * sx - cz * cx) * py + (-sz * sy * cx + cz * sx) * pz;
                j[8] = (cz * cy) * px + (cz * sy * sx - sz * cx) * py + (cz * sy * cx + sz * sx) * pz;
            } else {
                j[0] = (-sz * cx * sy) * px + (sz * sx) * py + (sz * cx * cy) * pz;
                j[6] = (cz * cx * sy) * px + (-cz * sx) * py + (-cz * cx * cy) * pz;
                j[12] = (sx * sy) * px + (cx) * py + (-sx * cy) * pz;
                j[1] = (-cz * sy - sz * sx * cy) * px + (cz * cy - sz * sx * sy) * pz;
                j[7] = (-sz * sy + cz * sx * cy) * px + (sz * cy + cz * sx * sy) * pz;
                j[13] = (-cx * cy) * px + (-cx * sy) * pz;
                j[2] = (-sz * cy - cz * sx * sy) * px + (-cz * cx) * py + (-sz * sy + cz * sx * cy) * pz;
                j[8] = (cz * cy - sz * sx * sy) * px + (-sz * cx) * py + (cz * sy + sz * sx * cy) * pz;
            }
            j[14] = 0.0;
            j[3] = 1.0;
            j[10] = 1.0;
            j[17] = 1.0;
        }
    }

    // The affine decomposition J(x) = J0 + sum_e x_e C_e of the Jacobian
    // (registration/cuda.rs:294-318): j0 = J(0), c[e] = J(e_e) - J(0);
    // c is laid out [e][d][k].
    void affine_form(double *j0, double *c) const {
        const double zero[3] = {0, 0, 0};
        jacobian(zero, j0);
        double je[36];
        for (int e = 0; e < 3; ++e) {
            double basis[3] = {0, 0, 0};
            basis[e] = 1.0;
            jacobian(basis, je);
            for (int i = 0; i < 3 * nparams; ++i) c[e * 3 * nparams + i] = je[i] - j0[i];
        }
    }
};

}  // namespace sitkhip
