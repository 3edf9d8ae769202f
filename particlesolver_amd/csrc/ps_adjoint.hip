// ps_adjoint.hip — reverse-mode adjoint of one substep (ps_step_backward).
//
// The forward map of a substep is predict -> I Jacobi iterations -> velocity
// + XSPH.  With the neighbour table, the contact sets and counts, the
// coordinates walls and pins set and the wall-friction branches frozen (as
// the recomputed forward pass recorded them), each piece is smooth and its
// vector-Jacobian product is written out by hand below.  Every transposed
// sum is a gather over the particle's own neighbour table, which holds the
// same pairs as its neighbours' tables (r^2 < R^2 is symmetric and no table
// overflowed), so there are no scatters and no atomics: one thread per
// particle (or per distance constraint), sums in table order.  The
// arithmetic is the CPU oracle's, operation by operation (built with
// -ffp-contract=off), so FP32 gradients match it bit for bit.
//
// Per iteration, with y = p + Delta before walls and pins:
//   Delta_i = f_i / rho0_i (fluid) + contacts + distance terms,
//   f_i = sum_j (l_i + l_j + s_ij) G_ij, G_ij = -b(r) d_ij,
//   b = spiky (h - r)^2 / r (G = (0, -gdeg, 0) for a degenerate pair),
//   l_i = -C_i / (D_i + eps), C_i = rho_i / rho0_i - 1,
//   D_i = sum_j a_ij^2 r^2 + |S_i|^2, S_i = sum_j a_ij d_ij, a = b / rho0_i.
// Kernels per reversed iteration: adj_lambda_k (dL/dl -> dL/dC, dL/dD),
// adj_position_k (dL/dp: pair, density, D, contact and distance terms, and
// dL/dstiffness per constraint), which ends with the next iteration's
// pre-step (adj_pre_one: walls, pins and wall friction backwards, the rho0
// term of f).  Per substep: adj_xsph_k and adj_finalize_k (velocity + XSPH
// backwards, then the last iteration's pre-step) before the iterations,
// adj_predict_k after them.

#include "ps_device.hpp"

namespace ps {

// The kernels that gather over the neighbour table map blocks with
// xcd_block(), so each XCD reverses one contiguous range of sorted particles
// and its L2 holds that region's records.
static constexpr uint32_t kAdjBlock = 256;

__device__ __forceinline__ float dot3(const float *a, const float *b)
{
    return ((a[0] * b[0]) + (a[1] * b[1])) + (a[2] * b[2]);
}

// G_ij for d = p_i - p_j (r2 = |d|^2 < h2)
__device__ __forceinline__ void adj_gamma(const KConst &k, const float *dv, float r2, float *g)
{
    if (r2 < kDegenerateR2) {
        g[0] = 0.0f;
        g[1] = -((k.spiky * (k.h * k.h)) * 0.001f);
        g[2] = 0.0f;
        return;
    }
    const float ri = rsq(r2);
    const float q = k.h - (r2 * ri);
    const float bb = (k.spiky * (q * q)) * ri;
    for (int a = 0; a < 3; a++) g[a] = -(bb * dv[a]);
}

// gradient of W_poly6(|d|) in d: -6 poly6 (h^2 - r^2)^2 d
__device__ __forceinline__ void adj_grad_w6(const KConst &k, const float *dv, float r2, float *g)
{
    const float t = k.h2 - r2;
    const float c = -((6.0f * k.poly6) * (t * t));
    for (int a = 0; a < 3; a++) g[a] = c * dv[a];
}

// Cloth contact of owner i with j (collideD without friction: a distance
// constraint of rest collide_dist, stiffness 1, count nc): M^T y for
// M = -(w_i / (wsum nc)) [(1 - cd/r) I + (cd/r^3) d d^T]; false when the
// pair is not in i's contact set.
__device__ __forceinline__ bool adj_contact(const KConst &k, float wi, int32_t ph, float wj, int32_t phj,
                                            uint32_t nci, const float *dv, const float *y, float *out)
{
    if (ph < PS_PHASE_CLOTH || wi == 0.0f || nci == 0u) return false;
    if (ph > PS_PHASE_SOLID && ph == phj) return false;
    const float r2 = ((dv[0] * dv[0]) + (dv[1] * dv[1])) + (dv[2] * dv[2]);
    if (!(r2 < k.collide_r2)) return false;
    const float dist = sqrtf(r2);
    if (dist < 1e-6f) return false;
    const float wsum = wi + wj;
    if (wsum == 0.0f) return false;
    const float cf = -((wi / wsum) / (float)nci);
    const float k1 = 1.0f - (k.collide_dist / dist), k2 = k.collide_dist / ((dist * dist) * dist);
    const float dy = dot3(dv, y);
    for (int a = 0; a < 3; a++) out[a] = cf * ((k1 * y[a]) + ((k2 * dy) * dv[a]));
    return true;
}

__device__ __forceinline__ void load3(const float4 &v, float *o)
{
    o[0] = v.x;
    o[1] = v.y;
    o[2] = v.z;
}

// masked dL/dy of particle i as the Jacobian uses it (z dropped in 2D)
__device__ __forceinline__ void yd_of(const KConst &k, const float4 &Y, float *yd)
{
    yd[0] = Y.x;
    yd[1] = Y.y;
    yd[2] = k.dims == 2u ? 0.0f : Y.z;
}

// Both owners' fluid terms of the pair (p, j) in dL/dp_p, d = p_p - p_j:
// E_p - E_j, with E_i = d/dd [mu_i . (l_i + l_j + s(d)) G(d)] for owner i
// (J_G = -(b I + (b'/r) d d^T), b' = -spiky (h - r)(h + r) / r^2), plus
// cm_p g6(d) - cm_j g6(-d) (the density term, g6 = dW/dd) and
// Db_p dD_p(d) - Db_j dD_j(-d) (dD_i/dd of the denominator D_i).  Folded
// into one expression,
//   (alpha - K mD + cms c6) d + 2A dV - cb dmu,
//   dmu = mu_p - mu_j, mD = d.dmu, K = b cc + coef b'/r, cb = coef b,
//   dV = V_p - V_j, alpha = C us - 2 Bq (d.dV), us = u_p + u_j,
//   A = (h - r)^2 / r, Bq = (h - r)(h + r) / r^3, C = -4 (h - r)^3 / r,
//   cms = Cr_p m_j + Cr_j m_p,
// with the per-particle factors of adj_lambda_k (MC: mu, Cr; KV: V, u).  An
// owner that does not take part (op, oj false) enters with zeros.  cc and
// c6 are the tensile term's and W's gradient factors (dW/dd = c6 d).  A
// degenerate pair (G the constant (0, g0y, 0) for both owners) adds
// (cc g0y (mu_p.y + mu_j.y) + cms c6) d.  Adds the terms to acc.
__device__ __forceinline__ void adj_pair_sum(const KConst &k, const float *dv, float r2, float lsum, bool op,
                                             bool oj, const float4 &mcp, const float4 &mcj, const float4 &kvp,
                                             const float4 &kvj, float massp, float massj, float *acc)
{
    const float t = k.h2 - r2;
    const float tt = t * t;
    const float W = k.poly6 * (tt * t);
    const float x = W * k.inv_w_dq;
    const uint32_t n = k.n_corr;
    const float s = -(k.k_corr * ipow(x, n));
    const float xn1 = n > 1u ? ipow(x, n - 1u) : 1.0f;
    const float cc = (((k.k_corr * (float)n) * xn1) * k.inv_w_dq) * ((6.0f * k.poly6) * tt);
    const float c6 = -((6.0f * k.poly6) * tt);
    const float coef = lsum + s;
    const float mup[3] = {op ? mcp.x : 0.0f, op ? mcp.y : 0.0f, op ? mcp.z : 0.0f};
    const float muj[3] = {oj ? mcj.x : 0.0f, oj ? mcj.y : 0.0f, oj ? mcj.z : 0.0f};
    const float cms = (op ? mcp.w * massj : 0.0f) + (oj ? mcj.w * massp : 0.0f);
    if (r2 < kDegenerateR2) {
        const float g0y = -((k.spiky * (k.h * k.h)) * 0.001f);
        const float be = (cc * (g0y * (mup[1] + muj[1]))) + (cms * c6);
        for (int a = 0; a < 3; a++) acc[a] = acc[a] + (be * dv[a]);
        return;
    }
    const float ri = rsq(r2);
    const float r = r2 * ri;
    const float q = k.h - r;
    const float qq = q * q;
    const float bb = (k.spiky * qq) * ri;
    const float bp = -(((k.spiky * q) * (k.h + r)) * (ri * ri));
    const float dmu[3] = {mup[0] - muj[0], mup[1] - muj[1], mup[2] - muj[2]};
    const float mD = dot3(dv, dmu);
    const float K = (bb * cc) + (coef * (bp * ri));
    const float cb = coef * bb;
    const float A2 = 2.0f * (qq * ri);
    const float Bq = ((q * (k.h + r)) * (ri * ri)) * ri;
    const float C = -((4.0f * (qq * q)) * ri);
    const float us = (op ? kvp.w : 0.0f) + (oj ? kvj.w : 0.0f);
    const float dV[3] = {(op ? kvp.x : 0.0f) - (oj ? kvj.x : 0.0f), (op ? kvp.y : 0.0f) - (oj ? kvj.y : 0.0f),
                         (op ? kvp.z : 0.0f) - (oj ? kvj.z : 0.0f)};
    const float vD = dot3(dv, dV);
    const float al = (C * us) - ((2.0f * Bq) * vD);
    const float be = (al - (K * mD)) + (cms * c6);
    for (int a = 0; a < 3; a++) acc[a] = acc[a] + ((be * dv[a]) + ((A2 * dV[a]) - (cb * dmu[a])));
}

// density multiplier of a neighbour in rho_i
__device__ __forceinline__ float adj_mass(const KConst &k, float wj, int32_t phj)
{
    float m = 1.0f / wj;
    if (phj != PS_PHASE_FLUID) m = m * k.s_solid;
    return m;
}

// Neighbour records of the gathers.  UNIFORM (every particle fluid, one
// inverse mass) variants leave out the attribute loads.
struct GLam {
    float4 p, m;
};
struct GLamAttr {
    float4 p, m;
    AdjAttr a;
};
struct GPos {
    float4 p, mc, kv;
};
struct GPosAttr {
    float4 p, mc, kv;
    AdjAttr a;
};
struct GXA {
    float4 p, x;
    Attr a;
};
struct GXY {
    float4 p, x, y;
};
struct GXYA {
    float4 p, x, y;
    Attr a;
};

// ---------------------------------------------------------------------------
// Pins, walls and wall friction backwards, and the fluid sum's rho0 term,
// of one particle (the pre-step of a reversed iteration, from the recorded
// forward quantities: f and the contact count of the solve, the clamp bits
// and wall-friction records of walls()), for g = dL/dp after the iteration.
// Wall friction (walls(): p'' = p' - t f, t = e - (e.n) n, e = p' - x*):
// static, or kinetic clamped to f = 1: dp''/dp' = n n^T,
// dp''/dx* = I - n n^T; kinetic with f = wall_k / l < 1:
// dp''/dp' = I - wall_k K (I - n n^T), K = (I - t^ t^T) / l, the x* part its
// negative second term, and dp''/dwall_k = -t / l.  Writes the iteration's
// Yn (masked dL/dy), M = (mu = yd / rho0 for fluid, yd otherwise; 1 / rho0)
// and the attribute record ABn the later kernels gather; updates xsb, wkb
// and r0b.  Run at the end of the kernel that forms g (adj_finalize_k,
// adj_position_k), by the particle's lane 0.
// The caller loads the records (AdjPreIn), early where registers allow, so
// that this tail waits on no load unless the particle has wall friction.
struct AdjPreIn {
    uint32_t cl;  // clamp bits
    float4 tf;    // recorded fluid sum and contact count
    float r0b;    // dL/drho0 so far
};

__device__ __forceinline__ AdjPreIn adj_pre_load(const AdjBuffers &a, uint32_t i)
{
    AdjPreIn q;
    q.cl = a.clampn[i];
    q.tf = a.tfn[i];
    q.r0b = a.r0b[i];
    return q;
}

template <bool UNIFORM>
__device__ __forceinline__ void adj_pre_one(const DevBuffers &b, const KConst &k, const AdjBuffers &a,
                                            uint32_t i, const Attr &ai, const AdjPreIn &in, float *g)
{
    const uint32_t cl = in.cl;
    const float4 tf = in.tf;
    if (cl & kPinned) {
        g[0] = g[1] = g[2] = 0.0f;
    } else if (cl & kWallFriction) {
        const float4 rec = a.wrecn[i];
        const float4 xs = b.X[i];
        const int code = (int)rec.w;
        const float nn[3] = {(float)(code / 9 - 1), (float)((code / 3) % 3 - 1), (float)(code % 3 - 1)};
        const float e[3] = {rec.x - xs.x, rec.y - xs.y, rec.z - xs.z};
        const float dn = dot3(e, nn);
        const float t[3] = {e[0] - (dn * nn[0]), e[1] - (dn * nn[1]), e[2] - (dn * nn[2])};
        float4 xb = a.xsb[i];
        float xv[3] = {xb.x, xb.y, xb.z};
        if (!(cl & kWallKinetic)) {
            const float ng = dot3(nn, g);
            for (int q = 0; q < 3; q++) {
                const float pa = ng * nn[q];
                xv[q] = xv[q] + (g[q] - pa);
                g[q] = pa;
            }
        } else {
            const float l = sqrtf(dot3(t, t));
            const float tg = dot3(t, g);
            float kg[3];
            for (int q = 0; q < 3; q++) kg[q] = (g[q] - ((t[q] * tg) / (l * l))) / l;
            const float nk = dot3(nn, kg);
            for (int q = 0; q < 3; q++) {
                const float qq = k.wall_k * (kg[q] - (nk * nn[q]));
                xv[q] = xv[q] + qq;
                g[q] = g[q] - qq;
            }
            a.wkb[i] = a.wkb[i] - (tg / l);
        }
        a.xsb[i] = make_float4(xv[0], xv[1], xv[2], 0.0f);
    }
    for (int q = 0; q < 3; q++)
        if (cl & (1u << q)) g[q] = 0.0f;
    const float yd[3] = {g[0], g[1], k.dims == 2u ? 0.0f : g[2]};
    const bool fl = UNIFORM || ai.phase == PS_PHASE_FLUID;
    const float ir0 = 1.0f / ai.rho0;
    if (fl) {
        const float f[3] = {tf.x, tf.y, tf.z};
        a.r0b[i] = in.r0b - (dot3(yd, f) * (ir0 * ir0));
    }
    a.Yn[i] = make_float4(g[0], g[1], g[2], 0.0f);
    a.M[i] = fl ? make_float4(yd[0] * ir0, yd[1] * ir0, yd[2] * ir0, ir0) : make_float4(yd[0], yd[1], yd[2], ir0);
    if (!UNIFORM) {
        const uint32_t nc = __float_as_uint(tf.w);
        a.nc[i] = nc;
        a.ABn[i] = AdjAttr{ai.w, ai.phase, nc, adj_mass(k, ai.w, ai.phase)};
    }
}

// dL/dl -> dL/dC = -lb / (D + eps), dL/dD = -lb l / (D + eps); rho0 terms.
// Four lanes per particle: lane hh sums lb over quarter hh of the table (its
// chunks 4m + hh), then lb = (q0 + q2) + (q1 + q3), and lane 0 finishes.
template <bool UNIFORM>
__global__ __launch_bounds__(kAdjBlock) void adj_lambda_k(DevBuffers b, KConst k, AdjBuffers a, uint32_t n)
{
    const uint32_t i = xcd_block() * (kAdjBlock / 4) + (threadIdx.x >> 2);
    const uint32_t hh = threadIdx.x & 3u;
    if (i >= n) return;  // the four lanes of a particle leave together
    // every per-particle load goes out before the first wait (lane 0's
    // finish would otherwise wait on three more after the sum)
    const float4 mi = a.M[i];
    const int32_t phi = UNIFORM ? (int32_t)PS_PHASE_FLUID : b.A[i].phase;
    const float4 pi = a.Pit[i];
    const uint32_t c = b.ncnt[i];
    const float4 sd = a.SD[i];
    const float rhoi = a.rho[i];
    float r = a.r0b[i];
    if (!UNIFORM && phi != PS_PHASE_FLUID) {
        if (!hh) {
            a.MC[i] = make_float4(mi.x, mi.y, mi.z, 0.0f);
            a.KV[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        }
        return;
    }
    const float mu[3] = {mi.x, mi.y, mi.z};
    const float ir0 = mi.w;
    float lb = 0.0f;
    auto body = [&](const float4 &pj, const float4 &mj, float wj, int32_t phj) {
        if (wj == 0.0f) return;
        const float dv[3] = {pi.x - pj.x, pi.y - pj.y, pi.z - pj.z};
        const float r2 = dist2(dv[0], dv[1], dv[2]);
        if (!(r2 < k.h2)) return;
        float g[3];
        adj_gamma(k, dv, r2, g);
        lb = lb + dot3(mu, g);
        if (phj == PS_PHASE_FLUID) {
            // G(-d) = -G(d) exactly (rounding is symmetric); the degenerate
            // direction is the same for both
            const bool deg = r2 < kDegenerateR2;
            const float gj[3] = {deg ? g[0] : -g[0], deg ? g[1] : -g[1], deg ? g[2] : -g[2]};
            const float muj[3] = {mj.x, mj.y, mj.z};
            lb = lb + dot3(muj, gj);
        }
    };
    if (UNIFORM) {
        for_neighbors_strided_chunked(
            b, i, c, hh, 4u, [&](uint32_t j) { return GLam{a.Pit[j], a.M[j]}; },
            [&](uint32_t, const GLam &d) { body(d.p, d.m, 1.0f, PS_PHASE_FLUID); });
    } else {
        for_neighbors_strided_chunked(
            b, i, c, hh, 4u, [&](uint32_t j) { return GLamAttr{a.Pit[j], a.M[j], a.AB[j]}; },
            [&](uint32_t, const GLamAttr &d) { body(d.p, d.m, d.a.w, d.a.phase); });
    }
    lb = lb + __shfl_xor(lb, 2, 64);  // lanes 0 and 1: q0 + q2, q1 + q3
    const float lo = __shfl_xor(lb, 1, 64);
    if (hh) return;
    lb = lb + lo;
    const float den = sd.w + k.relax;
    const float lam = pi.w;
    const float Cb = -(lb / den);
    const float Db = -((lb * lam) / den);
    r = r - (Cb * (rhoi * (ir0 * ir0)));
    r = r - (Db * ((2.0f * sd.w) * ir0));
    a.r0b[i] = r;
    // the factors adj_position_k combines per pair
    const float sk = k.spiky * ir0;
    const float dk = Db * sk;
    a.MC[i] = make_float4(mu[0], mu[1], mu[2], Cb * ir0);
    a.KV[i] = make_float4(dk * sd.x, dk * sd.y, dk * sd.z, Db * (sk * sk));
}

// dL/dp before the iteration.  Four lanes per particle: lane hh sums the
// pair and contact terms of quarter hh of the table (chunks hh, hh + 4, ...;
// quarter 0 starts from dL/dy), then acc = (q0 + q2) + (q1 + q3), and lane 0
// adds the distance constraints.
template <bool UNIFORM>
__global__ __launch_bounds__(kAdjBlock) void adj_position_k(DevBuffers b, KConst k, AdjBuffers a, uint32_t n)
{
    const uint32_t p = xcd_block() * (kAdjBlock / 4) + (threadIdx.x >> 2);
    const uint32_t hh = threadIdx.x & 3u;
    if (p >= n) return;  // the four lanes of a particle leave together
    const float4 pp = a.Pit[p];
    const Attr ap = b.A[p];
    const float4 Yp = a.Y[p];
    float acc[3] = {hh ? 0.0f : Yp.x, hh ? 0.0f : Yp.y, hh ? 0.0f : Yp.z};
    float ydp[3];
    yd_of(k, Yp, ydp);
    const bool pf = UNIFORM || ap.phase == PS_PHASE_FLUID;
    const float4 mcp = a.MC[p];  // mu_p, Cr_p
    const float4 kvp = a.KV[p];  // V_p, u_p
    const uint32_t ncp = UNIFORM ? 0u : a.nc[p];
    const bool contacts = !UNIFORM && a.contacts;
    const float massp = adj_mass(k, ap.w, ap.phase);  // p as a neighbour of j
    auto body = [&](const float4 &pj, const float4 &mcj, const float4 &kvj, float wj, int32_t phj, uint32_t ncj,
                    float massj) {
        const float dv[3] = {pp.x - pj.x, pp.y - pj.y, pp.z - pj.z};
        const float r2 = dist2(dv[0], dv[1], dv[2]);
        if (r2 < k.h2) {
            // p owns the pair when fluid and j moves, j when fluid and p moves
            const bool op = pf && wj != 0.0f;
            const bool oj = (UNIFORM || phj == PS_PHASE_FLUID) && ap.w != 0.0f;
            if (op || oj) adj_pair_sum(k, dv, r2, pp.w + pj.w, op, oj, mcp, mcj, kvp, kvj, massp, massj, acc);
        }
        if (contacts) {
            // cloth contacts: p's own, then p as the partner of j's (for a
            // non-fluid j, MC holds its masked dL/dy)
            const float dj[3] = {-dv[0], -dv[1], -dv[2]};
            float m[3];
            if (adj_contact(k, ap.w, ap.phase, wj, phj, ncp, dv, ydp, m))
                for (int q2 = 0; q2 < 3; q2++) acc[q2] = acc[q2] + m[q2];
            const float ydj[3] = {mcj.x, mcj.y, mcj.z};
            if (adj_contact(k, wj, phj, ap.w, ap.phase, ncj, dj, ydj, m))
                for (int q2 = 0; q2 < 3; q2++) acc[q2] = acc[q2] - m[q2];
        }
    };
    const uint32_t c = b.ncnt[p];
    // two records (48 or 64 bytes each) in flight per lane: four would hold
    // the kernel to four wavefronts per SIMD
    if (UNIFORM) {
        for_neighbors_strided_pairs(
            b, p, c, hh, 4u, [&](uint32_t j) { return GPos{a.Pit[j], a.MC[j], a.KV[j]}; },
            [&](uint32_t, const GPos &d) { body(d.p, d.mc, d.kv, ap.w, PS_PHASE_FLUID, 0u, massp); });
    } else {
        for_neighbors_strided_pairs(
            b, p, c, hh, 4u, [&](uint32_t j) { return GPosAttr{a.Pit[j], a.MC[j], a.KV[j], a.AB[j]}; },
            [&](uint32_t, const GPosAttr &d) { body(d.p, d.mc, d.kv, d.a.w, d.a.phase, d.a.nc, d.a.mass); });
    }
    {
        // lanes 0 and 1 form q0 + q2 and q1 + q3, lane 0 their sum
        for (int q2 = 0; q2 < 3; q2++) acc[q2] = acc[q2] + __shfl_xor(acc[q2], 2, 64);
        float lo[3];
        for (int q2 = 0; q2 < 3; q2++) lo[q2] = __shfl_xor(acc[q2], 1, 64);
        for (int q2 = 0; q2 < 3; q2++) acc[q2] = acc[q2] + lo[q2];  // used in lane 0
    }
    // distance constraints incident to p, in their order; for a constraint
    // whose first particle p is, also its dL/dstiffness (one writer per
    // constraint): T_a / st = -(w_a / (wsum occ_a)) (1 - L/r) d_ab, and T_b
    // its mirror.  The particle's four lanes take four constraints at a
    // time, kR such rounds at once: each constraint is a chain of dependent
    // loads (index, pair, slot, records), and each level of the chain goes
    // out for all kR rounds together.  Lane 0 adds the terms in constraint
    // order.
    if (a.dist) {
        constexpr uint32_t kR = 3;
        const uint32_t o = ap.orig;
        const uint32_t e0 = b.dist_off[o], e1 = b.dist_off[o + 1];
        const uint32_t quad = threadIdx.x & ~3u;
        const float occo = (float)b.occ[o];
        for (uint32_t base = e0; base < e1; base += 4u * kR) {  // the same trip count in the four lanes
            uint32_t cid[kR], j[kR];
            bool ok[kR];
            uint2 pr[kR];
#pragma unroll
            for (uint32_t r = 0; r < kR; r++) {
                const uint32_t e = base + 4u * r + hh;
                ok[r] = e < e1;
                cid[r] = b.dist_adj[ok[r] ? e : e0];
            }
#pragma unroll
            for (uint32_t r = 0; r < kR; r++) pr[r] = b.dist_pairs[cid[r]];
#pragma unroll
            for (uint32_t r = 0; r < kR; r++) j[r] = b.slot_of[pr[r].x == o ? pr[r].y : pr[r].x];
            float4 pj[kR], yj[kR];
            float2 rs[kR];
            Attr aj[kR];
#pragma unroll
            for (uint32_t r = 0; r < kR; r++) {
                pj[r] = a.Pit[j[r]];
                aj[r] = b.A[j[r]];
                yj[r] = a.Y[j[r]];
                rs[r] = b.dist_rs[cid[r]];
            }
            // each round's p term X and j term Z; fl bit 0: X is added,
            // bit 1: Z is subtracted
            float X[kR][3], Z[kR][3];
            uint32_t fl[kR];
#pragma unroll
            for (uint32_t r = 0; r < kR; r++) {
                fl[r] = 0;
                for (int q2 = 0; q2 < 3; q2++) X[r][q2] = Z[r][q2] = 0.0f;
                const float dv[3] = {pp.x - pj[r].x, pp.y - pj[r].y, pp.z - pj[r].z};
                const float dist = sqrtf(((dv[0] * dv[0]) + (dv[1] * dv[1])) + (dv[2] * dv[2]));
                const float wsum = ap.w + aj[r].w;
                if (!ok[r] || dist < 1e-6f || wsum == 0.0f) continue;
                const float L = rs[r].x, st = rs[r].y;
                const float k1 = 1.0f - (L / dist), k2 = L / ((dist * dist) * dist);
                const bool first = pr[r].x == o;  // d = p_a - p_b
                float gs = 0.0f;
                // M(d) = -(w st / (wsum occ)) [(1 - L/r) I + (L/r^3) d d^T]
                if (ap.w != 0.0f) {
                    const float cf = -(((ap.w * st) / wsum) / occo);
                    const float dy = dot3(dv, ydp);
                    for (int q2 = 0; q2 < 3; q2++) X[r][q2] = cf * ((k1 * ydp[q2]) + ((k2 * dy) * dv[q2]));
                    fl[r] |= 1u;
                    if (first) gs = gs - (((ap.w / wsum) / occo) * (k1 * dy));
                }
                if (aj[r].w != 0.0f) {
                    float ydj[3];
                    yd_of(k, yj[r], ydj);
                    const float occj = (float)b.occ[aj[r].orig];
                    const float cf = -(((aj[r].w * st) / wsum) / occj);
                    const float dy = dot3(dv, ydj);
                    for (int q2 = 0; q2 < 3; q2++) Z[r][q2] = cf * ((k1 * ydj[q2]) + ((k2 * dy) * dv[q2]));
                    fl[r] |= 2u;
                    if (first) gs = gs + (((aj[r].w / wsum) / occj) * (k1 * dy));
                }
                if (first) a.sb[cid[r]] = a.sb[cid[r]] + gs;
            }
#pragma unroll
            for (uint32_t r = 0; r < kR; r++)
                for (uint32_t l = 0; l < 4u; l++) {
                    const int src = (int)(quad + l);
                    const uint32_t fll = (uint32_t)__shfl((int)fl[r], src, 64);
                    float Xl[3], Zl[3];
                    for (int q2 = 0; q2 < 3; q2++) {
                        Xl[q2] = __shfl(X[r][q2], src, 64);
                        Zl[q2] = __shfl(Z[r][q2], src, 64);
                    }
                    if (fll & 1u)
                        for (int q2 = 0; q2 < 3; q2++) acc[q2] = acc[q2] + Xl[q2];
                    if (fll & 2u)
                        for (int q2 = 0; q2 < 3; q2++) acc[q2] = acc[q2] - Zl[q2];
                }
        }
    }
    if (hh) return;
    // the next iteration's pre-step (its records loaded here: held from the
    // start they would cost this kernel a wavefront per SIMD)
    if (a.clampn)
        adj_pre_one<UNIFORM>(b, k, a, p, ap, adj_pre_load(a, p), acc);
    else
        a.pb[p] = make_float4(acc[0], acc[1], acc[2], 0.0f);
}

// ---------------------------------------------------------------------------
// velocity + XSPH backwards: v_i = u_i/dt + ((c/rho0_i) S_i)/dt,
// S_i = sum_fluid j (u_j - u_i) m_j W_ij, u = p - x*.  adj_xsph_k forms
// dL/dS_i = (c/rho0_i) vbar_i / dt (in Y) and the rho0 term; adj_finalize_k
// gathers dL/dp (the final iterate) and dL/dx*.

template <bool UNIFORM>
__global__ __launch_bounds__(kAdjBlock) void adj_xsph_k(DevBuffers b, KConst k, AdjBuffers a, uint32_t n,
                                                     float dt)
{
    // four lanes per particle: lane hh forms quarter hh of S, lane 0 finishes
    const uint32_t i = xcd_block() * (kAdjBlock / 4) + (threadIdx.x >> 2);
    const uint32_t hh = threadIdx.x & 3u;
    if (i >= n) return;  // the four lanes of a particle leave together
    // the particle's own loads go out together
    const Attr ai = b.A[i];
    const float4 pi = b.P[i], xi = b.X[i];
    const uint32_t c = b.ncnt[i];
    float r0 = 0.0f;
    float Sh[3] = {0.0f, 0.0f, 0.0f};
    if (k.xsph_c != 0.0f && (UNIFORM || ai.phase == PS_PHASE_FLUID)) {
        const float ui[3] = {pi.x - xi.x, pi.y - xi.y, pi.z - xi.z};
        const float4 vb = a.gv[ai.orig];
        const float vbn[3] = {vb.x, vb.y, vb.z};
        const float cr = k.xsph_c / ai.rho0;
        // S exactly as finalize_k forms it: four quarters of the table
        float Sm[3] = {0.0f, 0.0f, 0.0f};
        auto body = [&](const float4 &pj, const float4 &xj, float wj) {
            const float r2 = dist2(pi.x - pj.x, pi.y - pj.y, pi.z - pj.z);
            if (!(r2 < k.h2)) return;
            const float t = k.h2 - r2;
            const float m = ((t * t) * t) * (1.0f / wj);
            Sm[0] = __builtin_fmaf((pj.x - xj.x) - ui[0], m, Sm[0]);
            Sm[1] = __builtin_fmaf((pj.y - xj.y) - ui[1], m, Sm[1]);
            Sm[2] = __builtin_fmaf((pj.z - xj.z) - ui[2], m, Sm[2]);
        };
        if (UNIFORM) {
            for_neighbors_strided_chunked(
                b, i, c, hh, 4u, [&](uint32_t j) { return PX{b.P[j], b.X[j]}; },
                [&](uint32_t, const PX &d) { body(d.p, d.x, ai.w); });
        } else {
            for_neighbors_strided_chunked(
                b, i, c, hh, 4u, [&](uint32_t j) { return GXA{b.P[j], b.X[j], b.A[j]}; },
                [&](uint32_t, const GXA &d) {
                    if (d.a.phase == PS_PHASE_FLUID) body(d.p, d.x, d.a.w);
                });
        }
        // (q0 + q2) + (q1 + q3)
        for (int q = 0; q < 3; q++) Sm[q] = Sm[q] + __shfl_xor(Sm[q], 2, 64);
        float lo[3];
        for (int q = 0; q < 3; q++) lo[q] = __shfl_xor(Sm[q], 1, 64);
        if (hh) return;
        const float S[3] = {k.poly6 * (Sm[0] + lo[0]), k.poly6 * (Sm[1] + lo[1]), k.poly6 * (Sm[2] + lo[2])};
        const float vS = dot3(vbn, S);
        r0 = r0 - (((k.xsph_c / (ai.rho0 * ai.rho0)) * vS) / dt);
        for (int q = 0; q < 3; q++) Sh[q] = (vbn[q] * cr) / dt;
    }
    if (hh) return;
    a.r0b[i] = r0;
    a.wkb[i] = 0.0f;
    a.Y[i] = make_float4(Sh[0], Sh[1], Sh[2], 0.0f);
}

// Four lanes per particle: lane hh sums the pair terms of quarter hh of the
// table (quarter 0 starts from dL/dx and dL/dv / dt), then
// (q0 + q2) + (q1 + q3).
template <bool UNIFORM>
__global__ __launch_bounds__(kAdjBlock) void adj_finalize_k(DevBuffers b, KConst k, AdjBuffers a, uint32_t n,
                                                         float dt)
{
    const uint32_t p = xcd_block() * (kAdjBlock / 4) + (threadIdx.x >> 2);
    const uint32_t hh = threadIdx.x & 3u;
    if (p >= n) return;  // the four lanes of a particle leave together
    // the particle's own loads go out together (and the last iteration's
    // pre-step records)
    const Attr ap = b.A[p];
    const float4 pp = b.P[p], xp = b.X[p];
    const float4 shp4 = a.Y[p];
    const uint32_t c = b.ncnt[p];
    AdjPreIn pre{};
    if (a.clampn && hh == 0) pre = adj_pre_load(a, p);
    const float4 vb = a.gv[ap.orig], xb = a.gx[ap.orig];
    float ub[3] = {vb.x / dt, vb.y / dt, vb.z / dt};
    float acc[3] = {xb.x, xb.y, xb.z};
    if (hh)
        for (int q = 0; q < 3; q++) ub[q] = acc[q] = 0.0f;
    if (k.xsph_c != 0.0f && (UNIFORM || ap.phase == PS_PHASE_FLUID)) {
        const float up[3] = {pp.x - xp.x, pp.y - xp.y, pp.z - xp.z};
        const float Shp[3] = {shp4.x, shp4.y, shp4.z};
        const float mp = 1.0f / ap.w;
        auto body = [&](const float4 &pj, const float4 &xj, const float4 &shj4, float wj) {
            const float dv[3] = {pp.x - pj.x, pp.y - pj.y, pp.z - pj.z};
            const float r2 = dist2(dv[0], dv[1], dv[2]);
            if (!(r2 < k.h2)) return;
            const float t = k.h2 - r2;
            const float W = k.poly6 * ((t * t) * t);
            float g6[3];
            adj_grad_w6(k, dv, r2, g6);
            const float uj[3] = {pj.x - xj.x, pj.y - xj.y, pj.z - xj.z};
            const float Shj[3] = {shj4.x, shj4.y, shj4.z};
            // p as owner: S_p += (u_j - u_p) m_j W; p as neighbour of j
            const float mj = 1.0f / wj;
            const float wb = mj * ((((uj[0] - up[0]) * Shp[0]) + ((uj[1] - up[1]) * Shp[1])) +
                                   ((uj[2] - up[2]) * Shp[2]));
            const float wb2 = mp * ((((up[0] - uj[0]) * Shj[0]) + ((up[1] - uj[1]) * Shj[1])) +
                                    ((up[2] - uj[2]) * Shj[2]));
            for (int q2 = 0; q2 < 3; q2++) {
                ub[q2] = (ub[q2] - ((mj * W) * Shp[q2])) + ((mp * W) * Shj[q2]);
                acc[q2] = (acc[q2] + (wb * g6[q2])) + (wb2 * g6[q2]);
            }
        };
        if (UNIFORM) {
            for_neighbors_strided_chunked(
                b, p, c, hh, 4u, [&](uint32_t j) { return GXY{b.P[j], b.X[j], a.Y[j]}; },
                [&](uint32_t, const GXY &d) { body(d.p, d.x, d.y, ap.w); });
        } else {
            for_neighbors_strided_chunked(
                b, p, c, hh, 4u, [&](uint32_t j) { return GXYA{b.P[j], b.X[j], a.Y[j], b.A[j]}; },
                [&](uint32_t, const GXYA &d) {
                    if (d.a.phase == PS_PHASE_FLUID) body(d.p, d.x, d.y, d.a.w);
                });
        }
    }
    // lanes 0 and 1 form q0 + q2 and q1 + q3, lane 0 their sum
    for (int q = 0; q < 3; q++) {
        ub[q] = ub[q] + __shfl_xor(ub[q], 2, 64);
        acc[q] = acc[q] + __shfl_xor(acc[q], 2, 64);
    }
    float lu[3], la[3];
    for (int q = 0; q < 3; q++) {
        lu[q] = __shfl_xor(ub[q], 1, 64);
        la[q] = __shfl_xor(acc[q], 1, 64);
    }
    if (hh) return;
    for (int q = 0; q < 3; q++) {
        ub[q] = ub[q] + lu[q];
        acc[q] = acc[q] + la[q];
    }
    a.xsb[p] = make_float4(-ub[0], -ub[1], -ub[2], 0.0f);
    float g[3] = {acc[0] + ub[0], acc[1] + ub[1], acc[2] + ub[2]};
    if (a.clampn)
        adj_pre_one<UNIFORM>(b, k, a, p, ap, pre, g);  // the last iteration's pre-step
    else
        a.pb[p] = make_float4(g[0], g[1], g[2], 0.0f);
}

// predict backwards: p0 = x* + dt (v + dt g) for w != 0 (z held in 2D);
// results to insertion order, rho0 and wall_k terms accumulated there
__global__ __launch_bounds__(kAdjBlock) void adj_predict_k(DevBuffers b, KConst k, AdjBuffers a, uint32_t n,
                                                        float dt)
{
    const uint32_t t = blockIdx.x * kAdjBlock + threadIdx.x;
    if (t >= n) return;
    const Attr at = b.A[t];
    const uint32_t o = at.orig;
    const float4 xs = a.xsb[t], pb = a.pb[t];
    const bool mv = at.w != 0.0f;
    const bool twod = k.dims == 2u;
    a.gx[o] = make_float4(xs.x + pb.x, xs.y + pb.y, xs.z + pb.z, 0.0f);
    a.gv[o] = make_float4(mv ? dt * pb.x : 0.0f, mv ? dt * pb.y : 0.0f, (mv && !twod) ? dt * pb.z : 0.0f,
                          0.0f);
    a.drho[o] = a.drho[o] + a.r0b[t];
    a.dwk[o] = a.dwk[o] + a.wkb[t];
}

// ---------------------------------------------------------------------------

static inline uint32_t adj_blocks(uint32_t n) { return (n + kAdjBlock - 1) / kAdjBlock; }

hipError_t launch_adj(hipStream_t st, AdjKernel which, const DevBuffers &b, const KConst &k,
                      const AdjBuffers &a, uint32_t count, float dt, bool uniform)
{
    if (count == 0) return hipSuccess;
    const dim3 grid(adj_blocks(count)), block(kAdjBlock);
    // four lanes per particle: kAdjBlock / 4 particles per block
    const dim3 grid4((count + kAdjBlock / 4 - 1) / (kAdjBlock / 4));
#define PS_ADJ_LAUNCH4(kern, ...)                                                                 \
    do {                                                                                           \
        if (uniform)                                                                               \
            PS_LAUNCH(kern<true>, grid4, block, 0, st, __VA_ARGS__);                      \
        else                                                                                       \
            PS_LAUNCH(kern<false>, grid4, block, 0, st, __VA_ARGS__);                     \
    } while (0)
    switch (which) {
    case ADJ_XSPH: PS_ADJ_LAUNCH4(adj_xsph_k, b, k, a, count, dt); break;
    case ADJ_FINALIZE: PS_ADJ_LAUNCH4(adj_finalize_k, b, k, a, count, dt); break;
    case ADJ_LAMBDA: PS_ADJ_LAUNCH4(adj_lambda_k, b, k, a, count); break;
    case ADJ_POSITION: PS_ADJ_LAUNCH4(adj_position_k, b, k, a, count); break;
    case ADJ_PREDICT: PS_LAUNCH(adj_predict_k, grid, block, 0, st, b, k, a, count, dt); break;
    default: return hipErrorInvalidValue;
    }
#undef PS_ADJ_LAUNCH4
    return hipGetLastError();
}

}  // namespace ps
