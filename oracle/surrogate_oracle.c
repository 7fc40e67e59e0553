/*
 * surrogate_oracle.c -- CPU restatement of the learned macro-step surrogate (TEST
 * INFRASTRUCTURE: the checker of the HIP forward, backward and Adam kernels; never linked
 * into the product).
 *
 * The network has no counterpart in the reference (its repository holds no learned
 * component); the definition restated here is the one the HIP path implements:
 *
 *   edge features of target i and neighbour j (j != i, d^2 <= R^2, the reference's
 *   neighbour set of src/spatial_hash.cpp:31-57), all FP32:
 *     r = x_i - x_j, d2 = (rx^2 + ry^2) + rz^2, |r| = sqrt(d2), q = |r| / h
 *     e = [r * inv_r (3), (v_i - v_j) * inv_v (3), |r| * inv_r, w_shape(q) * w_norm,
 *          (rho_j - rho0) * rho_scale]
 *   message:  u_ij = W1 e_ij + b1 (h1),  S_i = msg_scale * sum_j ReLU(u_ij)
 *             m_i  = W2 S_i + msg_scale * n_i * b2 (h2; n_i = number of edges)
 *   node:     x_i = [m_i, (rho_i - rho0) * rho_scale, v_i * inv_v]  (h2 + 4)
 *             z_i = ReLU(Wn1 x_i + bn1) (h3),  out_i = Wn2 z_i + bn2 (3)
 *   predicted macro-step acceleration  a_i = acc_scale * out_i
 *   loss      L = 1/(3B) sum_{i in batch} |out_i - t_i|^2   (t = target / acc_scale)
 *
 * The second message layer is linear, so sum_j (W2 h_ij + b2) = W2 sum_j h_ij + n_i b2
 * and it is applied once per particle.  Parameters are one flat FP32 array:
 *   W1[h1][9] b1[h1] W2[h2][h1] b2[h2] Wn1[h3][h2+4] bn1[h3] Wn2[3][h3] bn2[3]
 * (row-major, output index first).  With cfg.bf16 set, values are rounded to bfloat16
 * (round to nearest even) where the device rounds them: every parameter, the edge
 * features, S_i, the node input x_i and z_i.  All other arithmetic is FP64.
 *
 * Pinning (tests/test_surrogate_oracle.py): central finite differences of the loss on a
 * small net (FP64, no rounding); the ReLU-sum identity and the hoisted second layer
 * against a direct per-edge evaluation; Adam against its closed form under a constant
 * gradient.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "sph_oracle.h"

#ifdef _OPENMP
#include <omp.h>
#endif

float or_bf16_round(float f) {
    uint32_t u;
    memcpy(&u, &f, 4);
    if ((u & 0x7f800000u) == 0x7f800000u && (u & 0x007fffffu)) {
        u |= 0x00400000u;  /* quiet NaN */
    } else {
        u += 0x7fffu + ((u >> 16) & 1u);
    }
    u &= 0xffff0000u;
    memcpy(&f, &u, 4);
    return f;
}

size_t or_sur_nparams(const or_sur_cfg* c) {
    const size_t h1 = (size_t)c->h1, h2 = (size_t)c->h2, h3 = (size_t)c->h3;
    return h1 * OR_SUR_NF + h1 + h2 * h1 + h2 + h3 * (h2 + 4) + h3 + 3 * h3 + 3;
}

typedef struct {
    size_t W1, b1, W2, b2, Wn1, bn1, Wn2, bn2;
} offs_t;

static offs_t offsets(const or_sur_cfg* c) {
    const size_t h1 = (size_t)c->h1, h2 = (size_t)c->h2, h3 = (size_t)c->h3;
    offs_t o;
    o.W1 = 0;
    o.b1 = o.W1 + h1 * OR_SUR_NF;
    o.W2 = o.b1 + h1;
    o.b2 = o.W2 + h2 * h1;
    o.Wn1 = o.b2 + h2;
    o.bn1 = o.Wn1 + h3 * (h2 + 4);
    o.Wn2 = o.bn1 + h3;
    o.bn2 = o.Wn2 + 3 * h3;
    return o;
}

static double rnd(const or_sur_cfg* c, double v) {
    return c->bf16 ? (double)or_bf16_round((float)v) : v;
}

/* Cubic-spline shape (CubicSplineKernel::W_3d without sigma, src/kernels.cpp:58-66). */
static float w_shape(float q) {
    if (q <= 1.0f) {
        float q2 = q * q;
        return (2.0f / 3.0f - q2) + 0.5f * q2 * q;
    }
    if (q <= 2.0f) {
        float t = 2.0f - q;
        return (1.0f / 6.0f) * t * t * t;
    }
    return 0.0f;
}

/* Edge features of (i, j) in FP32, rounded per cfg. */
static void edge_features(const or_sur_cfg* c, const float* pos, const float* vel,
                          const float* rho, int64_t i, int64_t j, double e[OR_SUR_NF]) {
    const float rx = pos[3 * i] - pos[3 * j];
    const float ry = pos[3 * i + 1] - pos[3 * j + 1];
    const float rz = pos[3 * i + 2] - pos[3 * j + 2];
    const float d2 = (rx * rx + ry * ry) + rz * rz;
    const float r = sqrtf(d2);
    const float q = r * c->inv_h;
    float f[OR_SUR_NF];
    f[0] = rx * c->inv_r;
    f[1] = ry * c->inv_r;
    f[2] = rz * c->inv_r;
    f[3] = (vel[3 * i] - vel[3 * j]) * c->inv_v;
    f[4] = (vel[3 * i + 1] - vel[3 * j + 1]) * c->inv_v;
    f[5] = (vel[3 * i + 2] - vel[3 * j + 2]) * c->inv_v;
    f[6] = r * c->inv_r;
    f[7] = w_shape(q) * c->w_norm;
    f[8] = (rho[j] - c->rho0) * c->rho_scale;
    for (int k = 0; k < OR_SUR_NF; ++k) e[k] = rnd(c, (double)f[k]);
}

/* Parameters as the device sees them. */
static double* load_params(const or_sur_cfg* c, const float* params) {
    const size_t np = or_sur_nparams(c);
    double* p = (double*)malloc(np * sizeof(double));
    for (size_t k = 0; k < np; ++k) p[k] = rnd(c, (double)params[k]);
    return p;
}

/* Per-particle forward state. */
typedef struct {
    double* S;    /* h1 values of S_i */
    double* x;    /* h2 + 4 */
    double* zp;   /* h3 pre-activation */
    double* z;    /* h3 */
    double out[3];
    int64_t n;
} node_t;

static void node_forward(const or_sur_cfg* c, const double* p, const offs_t* o,
                         const float* pos, const float* vel, const float* rho,
                         const int64_t* off, const int64_t* idx, int64_t i, node_t* nd) {
    const int h1 = c->h1, h2 = c->h2, h3 = c->h3;
    double e[OR_SUR_NF];
    for (int a = 0; a < h1; ++a) nd->S[a] = 0.0;
    nd->n = 0;
    for (int64_t t = off[i]; t < off[i + 1]; ++t) {
        const int64_t j = idx[t];
        if (j == i) continue;
        edge_features(c, pos, vel, rho, i, j, e);
        for (int a = 0; a < h1; ++a) {
            double u = p[o->b1 + a];
            for (int k = 0; k < OR_SUR_NF; ++k) u += p[o->W1 + (size_t)a * OR_SUR_NF + k] * e[k];
            if (u > 0.0) nd->S[a] += u;
        }
        nd->n++;
    }
    for (int a = 0; a < h1; ++a) nd->S[a] = rnd(c, (double)c->msg_scale * nd->S[a]);
    const double cn = (double)c->msg_scale * (double)nd->n;
    for (int b = 0; b < h2; ++b) {
        double m = cn * p[o->b2 + b];
        for (int a = 0; a < h1; ++a) m += p[o->W2 + (size_t)b * h1 + a] * nd->S[a];
        nd->x[b] = rnd(c, m);
    }
    nd->x[h2] = rnd(c, (double)((rho[i] - c->rho0) * c->rho_scale));
    for (int k = 0; k < 3; ++k) nd->x[h2 + 1 + k] = rnd(c, (double)(vel[3 * i + k] * c->inv_v));
    const int nin = h2 + 4;
    for (int b = 0; b < h3; ++b) {
        double zp = p[o->bn1 + b];
        for (int a = 0; a < nin; ++a) zp += p[o->Wn1 + (size_t)b * nin + a] * nd->x[a];
        nd->zp[b] = zp;
        nd->z[b] = zp > 0.0 ? rnd(c, zp) : 0.0;
    }
    for (int k = 0; k < 3; ++k) {
        double y = p[o->bn2 + k];
        for (int b = 0; b < h3; ++b) y += p[o->Wn2 + (size_t)k * h3 + b] * nd->z[b];
        nd->out[k] = y;
    }
}

static void node_alloc(const or_sur_cfg* c, node_t* nd) {
    nd->S = (double*)malloc((size_t)c->h1 * sizeof(double));
    nd->x = (double*)malloc((size_t)(c->h2 + 4) * sizeof(double));
    nd->zp = (double*)malloc((size_t)c->h3 * sizeof(double));
    nd->z = (double*)malloc((size_t)c->h3 * sizeof(double));
}

static void node_free(node_t* nd) {
    free(nd->S);
    free(nd->x);
    free(nd->zp);
    free(nd->z);
}

void or_sur_forward(const or_sur_cfg* c, const float* params, size_t n, const float* pos,
                    const float* vel, const float* rho, const int64_t* off, const int64_t* idx,
                    double* out, double* S_out, int64_t* count_out) {
    double* p = load_params(c, params);
    const offs_t o = offsets(c);
#pragma omp parallel
    {
        node_t nd;
        node_alloc(c, &nd);
#pragma omp for schedule(dynamic, 64)
        for (int64_t i = 0; i < (int64_t)n; ++i) {
            node_forward(c, p, &o, pos, vel, rho, off, idx, i, &nd);
            for (int k = 0; k < 3; ++k) out[3 * i + k] = nd.out[k];
            if (S_out)
                for (int a = 0; a < c->h1; ++a) S_out[(size_t)i * c->h1 + a] = nd.S[a];
            if (count_out) count_out[i] = nd.n;
        }
        node_free(&nd);
    }
    free(p);
}

double or_sur_loss_grad(const or_sur_cfg* c, const float* params, size_t n, const float* pos,
                        const float* vel, const float* rho, const int64_t* off,
                        const int64_t* idx, const uint8_t* in_batch, const float* target,
                        double* grad) {
    const size_t np = or_sur_nparams(c);
    double* p = load_params(c, params);
    const offs_t o = offsets(c);
    const int h1 = c->h1, h2 = c->h2, h3 = c->h3, nin = h2 + 4;
    int64_t B = 0;
    for (size_t i = 0; i < n; ++i) B += in_batch[i] ? 1 : 0;
    memset(grad, 0, np * sizeof(double));
    double loss = 0.0;
    if (B == 0) {
        free(p);
        return 0.0;
    }
    const double scale = 2.0 / (3.0 * (double)B);
#pragma omp parallel reduction(+ : loss)
    {
        node_t nd;
        node_alloc(c, &nd);
        double* g = (double*)calloc(np, sizeof(double));
        double* dz = (double*)malloc((size_t)h3 * sizeof(double));
        double* dm = (double*)malloc((size_t)h2 * sizeof(double));
        double* dS = (double*)malloc((size_t)h1 * sizeof(double));
        double e[OR_SUR_NF];
#pragma omp for schedule(dynamic, 16)
        for (int64_t i = 0; i < (int64_t)n; ++i) {
            if (!in_batch[i]) continue;
            node_forward(c, p, &o, pos, vel, rho, off, idx, i, &nd);
            double dout[3];
            for (int k = 0; k < 3; ++k) {
                const double r = nd.out[k] - (double)target[3 * i + k];
                loss += r * r;
                dout[k] = scale * r;
                g[o.bn2 + k] += dout[k];
                for (int b = 0; b < h3; ++b) g[o.Wn2 + (size_t)k * h3 + b] += dout[k] * nd.z[b];
            }
            for (int b = 0; b < h3; ++b) {
                double s = 0.0;
                for (int k = 0; k < 3; ++k) s += p[o.Wn2 + (size_t)k * h3 + b] * dout[k];
                dz[b] = nd.zp[b] > 0.0 ? s : 0.0;
                g[o.bn1 + b] += dz[b];
                for (int a = 0; a < nin; ++a) g[o.Wn1 + (size_t)b * nin + a] += dz[b] * nd.x[a];
            }
            for (int a = 0; a < h2; ++a) {
                double s = 0.0;
                for (int b = 0; b < h3; ++b) s += p[o.Wn1 + (size_t)b * nin + a] * dz[b];
                dm[a] = s;
            }
            const double cn = (double)c->msg_scale * (double)nd.n;
            for (int b = 0; b < h2; ++b) {
                g[o.b2 + b] += cn * dm[b];
                for (int a = 0; a < h1; ++a) g[o.W2 + (size_t)b * h1 + a] += dm[b] * nd.S[a];
            }
            for (int a = 0; a < h1; ++a) {
                double s = 0.0;
                for (int b = 0; b < h2; ++b) s += p[o.W2 + (size_t)b * h1 + a] * dm[b];
                dS[a] = (double)c->msg_scale * s;
            }
            for (int64_t t = off[i]; t < off[i + 1]; ++t) {
                const int64_t j = idx[t];
                if (j == i) continue;
                edge_features(c, pos, vel, rho, i, j, e);
                for (int a = 0; a < h1; ++a) {
                    double u = p[o.b1 + a];
                    for (int k = 0; k < OR_SUR_NF; ++k)
                        u += p[o.W1 + (size_t)a * OR_SUR_NF + k] * e[k];
                    if (u > 0.0) {
                        g[o.b1 + a] += dS[a];
                        for (int k = 0; k < OR_SUR_NF; ++k)
                            g[o.W1 + (size_t)a * OR_SUR_NF + k] += dS[a] * e[k];
                    }
                }
            }
        }
#pragma omp critical
        for (size_t k = 0; k < np; ++k) grad[k] += g[k];
        free(g);
        free(dz);
        free(dm);
        free(dS);
        node_free(&nd);
    }
    free(p);
    return loss / (3.0 * (double)B);
}

void or_sur_adam(size_t np, float* w, float* m, float* v, const float* g, float lr, float beta1,
                 float beta2, float eps, int32_t t) {
    const double bc1 = 1.0 - pow((double)beta1, (double)t);
    const double bc2 = 1.0 - pow((double)beta2, (double)t);
    for (size_t k = 0; k < np; ++k) {
        const double mk = (double)beta1 * m[k] + (1.0 - (double)beta1) * g[k];
        const double vk = (double)beta2 * v[k] + (1.0 - (double)beta2) * (double)g[k] * g[k];
        m[k] = (float)mk;
        v[k] = (float)vk;
        const double mh = mk / bc1, vh = vk / bc2;
        w[k] = (float)((double)w[k] - (double)lr * mh / (sqrt(vh) + (double)eps));
    }
}
