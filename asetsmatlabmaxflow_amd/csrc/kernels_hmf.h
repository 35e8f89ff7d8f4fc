// kernels_hmf.h -- hierarchical max-flow on a label tree, full-flow variant
// (maxflow/asetsHMF2D.m, asetsHMF3D.m: MaxFullFlow).  One iteration is
//   k_hmf_sink    sink flows of every node, bottom-up, one thread per voxel walking
//                 the whole tree (UpdateSinkFlows, asetsHMF3D.m:307-336)
//   kf_hmf_flow   labels and spatial flows of every non-root node (UpdateLabels
//                 :191-202, UpdateSpatialFlows :205-304): one z-streaming kernel of
//                 dependency radius 2, laid out as the binary model's fused kernel
//   kh_* passes   the same step as separate whole-field passes: the general path,
//                 and the only one for nodes with a shape constraint (its exemption
//                 step widens the dependency radius to 3)
//
// Arithmetic follows the reference's whole-array statements element by element in
// fp32 (the class of single-precision Ct), with:
//   g    = steps*(((div + pt) - P.pt) - u/cc)      u already updated this iteration
//   face = (face + g_hi) - g_lo  (3D, :210-212)    face + (g_hi - g_lo)  (2D, :225-226)
//   |p|  = sqrt(((((y_lo^2 + y_hi^2) + x_hi^2) + x_lo^2) + z_hi^2) + z_lo^2)
//   cap  = |p| <= alpha ? 1 : alpha/|p|            (correctly rounded sqrt and /)
//   face = [ex +] (0.5*face)*(cap_hi + cap_lo)
//   div  = ((((y_hi - y_lo) + x_hi) - x_lo) + z_hi) - z_lo
// where y is dim 1 (the reference's px, stride 1), x dim 2 (py), z dim 3 (pz), and a
// face outside the volume reads 0.  Faces are stored own-face (the face between
// voxel g - stride and g at g; axis index 0 holds the zero boundary face).
#pragma once
#include "common.h"
#include "kernels_fused.h"

namespace asets {

// One node of the tree, as the kernels see it.
struct HmfNode {
    const float *Ct;      // data cost of a leaf, else nullptr
    const float *alpha;   // per-voxel alpha, or nullptr: alpha_s
    float alpha_s;
    int32_t parent;       // -1 for the root
    int32_t shape;        // slot of the shape-constraint directions, -1: none
    int32_t pad;
};

// Sink-flow program: the tree in depth-first order, one event per visit.
enum : int32_t { kHmfEnter = 0, kHmfLeaf = 1, kHmfExit = 2 };
struct HmfEvent {
    int32_t type, node, depth;
    float count;   // divisor of an internal node's pt: children (+1 below the root)
    int32_t run;   // leaf: consecutive leaf events of the same parent from this one
};

// Sink flows of every node at one voxel.  The reference recurses children first;
// an internal node's pt = (base + sum over children, in order, of
// (pt_c + div_c) - u_c/cc) / count, where base = 1/cc at the root and
// (P.pt - div) + u/cc elsewhere, P.pt being the parent's value of the previous
// iteration (the parent is updated after its children); a leaf's pt =
// min(Ct, (P.pt - div) + u/cc).  Each thread keeps, per open ancestor depth, the
// running sum, the node's previous pt and its div and u/cc in LDS.
__global__ __launch_bounds__(256) void k_hmf_sink(const HmfNode *__restrict__ nodes,
                                                  const HmfEvent *__restrict__ ev, int nev,
                                                  const float *__restrict__ u, float *pt,
                                                  const float *__restrict__ div, int64_t N,
                                                  float cc, int fast_cc, float one_cc) {
    extern __shared__ float stk[];   // [depth][4][blockDim]
    const int T = blockDim.x, t = threadIdx.x;
    const float rcc = rcp_refined(cc);
    const int64_t g = (int64_t)blockIdx.x * T + t;
    if (g >= N) return;
    auto S = [&](int d, int j) -> float & { return stk[(d * 4 + j) * T + t]; };
    for (int e = 0; e < nev; ++e) {
        const HmfEvent E = ev[e];
        const int64_t i = (int64_t)E.node * N + g;
        if (E.type == kHmfEnter) {
            float acc = one_cc, d = 0.0f, uc = 0.0f;
            const float po = pt[i];
            if (E.depth > 0) {
                d = div[i];
                uc = div_fast(u[i], cc, rcc, fast_cc != 0);
                acc = (S(E.depth - 1, 1) - d) + uc;
            }
            S(E.depth, 0) = acc;
            S(E.depth, 1) = po;
            S(E.depth, 2) = d;
            S(E.depth, 3) = uc;
        } else if (E.type == kHmfLeaf && E.run >= 4) {
            // four sibling leaves: their loads issued together, then the sink
            // flows and the parent's running sum in child order
            float d[4], uu[4], ct[4];
            int64_t ii[4];
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int n = ev[e + j].node;
                ii[j] = (int64_t)n * N + g;
                d[j] = div[ii[j]];
                uu[j] = u[ii[j]];
                ct[j] = nodes[n].Ct[g];
            }
            const float po = S(E.depth - 1, 1);
            float acc = S(E.depth - 1, 0);
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const float uc = div_fast(uu[j], cc, rcc, fast_cc != 0);
                const float f = (po - d[j]) + uc;
                const float p = (ct[j] < f) ? ct[j] : f;
                pt[ii[j]] = p;
                acc = ((acc + p) + d[j]) - uc;
            }
            S(E.depth - 1, 0) = acc;
            e += 3;
        } else if (E.type == kHmfLeaf) {
            const float d = div[i];
            const float uc = div_fast(u[i], cc, rcc, fast_cc != 0);
            const float f = (S(E.depth - 1, 1) - d) + uc;
            const float ct = nodes[E.node].Ct[g];
            const float p = (ct < f) ? ct : f;
            pt[i] = p;
            S(E.depth - 1, 0) = ((S(E.depth - 1, 0) + p) + d) - uc;
        } else {
            const float p = S(E.depth, 0) / E.count;
            pt[i] = p;
            if (E.depth > 0)
                S(E.depth - 1, 0) = ((S(E.depth - 1, 0) + p) + S(E.depth, 2)) - S(E.depth, 3);
        }
    }
}

// ---- fused labels + spatial flows ------------------------------------------
struct HmfArgs {
    const HmfNode *nodes;
    const int32_t *flow;      // [nflow] indices of the non-root nodes
    int nflow;
    const float *u_in;        // [M][N]; u_out gets the updated labels
    float *u_out;
    const float *pt;          // [M][N]
    float *div;               // [M][N], written
    const float *f_in[3];     // own faces along y (dim 1), x (dim 2), z (dim 3)
    float *f_out[3];
    int64_t N;
    int ny, nx, nz, plane;
    float cc, steps;
    int fast_cc;
    int zchunk, ntile_y, ntile_x, nzc;
    int nblocks;              // nflow * ntile_y * ntile_x * nzc
};

__device__ __forceinline__ float hmf_cap(float t, float al) {
    const float m = __builtin_sqrtf(t);
    return (m <= al) ? 1.0f : al / m;
}

// Tile pipeline (the binary kernel's, kernels_fused.h): a workgroup owns a 64 x TX
// tile of one node and a range of z planes.  Per plane step, phase X: output of
// plane s-1, g of plane s+2 with its stepped z face; phase Y: stepped in-plane
// faces of s+2, capacity factors of s+1.  The updated labels of the tile are
// stored when their g is formed.
template <bool IS3D, int TX, int NT>
__device__ __forceinline__ void hmf_tile(const HmfArgs &a, int lb) {
    // streaming stores in 3D (kStream); the 2D tree's fields stay in the caches
    // between the flow and sink launches, where streaming measured 15 % slower
    constexpr int CP = IS3D ? kStream : 0;
    constexpr int TY = 64;
    constexpr int W2 = TY + 4, H2 = TX + 4;
    constexpr int W1 = TY + 2, H1 = TX + 2;
    constexpr int N2 = W2 * H2, N1 = W1 * H1;
    constexpr int K2 = (N2 + NT - 1) / NT;
    constexpr int K1 = (N1 + NT - 1) / NT;
    constexpr int KO = (TY * TX) / NT;
    constexpr int NH = N2 - TY * TX, NH1 = N1 - TY * TX;   // halo positions
    static_assert((TY * TX) % NT == 0, "tile must be a multiple of the block");
    static_assert(KO + (NH + NT - 1) / NT == K2, "radius-2 slots: own positions, then halo");
    static_assert(KO + (NH1 + NT - 1) / NT == K1, "radius-1 slots: own positions, then halo");
    static_assert(K2 <= 8, "flag word holds 8 positions");
    static_assert(K1 <= 32, "flag word holds 32 positions");

    __shared__ float GK[2][N2];
    __shared__ float B1X[3][N2], B1Y[3][N2], B1Z[IS3D ? 3 : 1][IS3D ? N2 : 1];
    __shared__ float G[2][N1];
    const int tid = threadIdx.x;
    // slot k holds a position (only the last halo slot can run past the halo)
    auto in2 = [&](int k) { return k < KO || (NH % NT) == 0 || tid + (k - KO) * NT < NH; };
    auto in1 = [&](int k) { return k < KO || (NH1 % NT) == 0 || tid + (k - KO) * NT < NH1; };
    // node slowest: a node's neighbouring tiles, which re-read each other's halo, run
    // together (measured 2.4 % faster at 192^3 than node fastest, where siblings
    // shared their parent's sink-flow halo instead)
    const int per_node = a.nblocks / a.nflow;
    const int slot = lb / per_node;
    const int rest = lb % per_node;
    const int n = a.flow[slot];
    const HmfNode nd = a.nodes[n];
    const int64_t on = (int64_t)n * a.N, op = (int64_t)nd.parent * a.N;
    const float *u_in = a.u_in + on, *pt = a.pt + on, *ppt = a.pt + op;
    float *u_out = a.u_out + on, *dvo = a.div + on;
    const float *fy = a.f_in[0] + on, *fx = a.f_in[1] + on;
    const float *fz = IS3D ? a.f_in[2] + on : nullptr;
    float *oy = a.f_out[0] + on, *ox = a.f_out[1] + on;
    float *oz = IS3D ? a.f_out[2] + on : nullptr;
    const float *amap = nd.alpha;
    const float as = nd.alpha_s;

    const int tyi = rest % a.ntile_y;
    const int txi = (rest / a.ntile_y) % a.ntile_x;
    const int zci = rest / (a.ntile_y * a.ntile_x);
    const int y0 = tyi * TY, x0 = txi * TX;
    const int z0 = zci * a.zchunk;
    const int z1 = min(z0 + a.zchunk, a.nz);
    const int ny = a.ny, nx = a.nx, nz = a.nz, plane = a.plane;
    const float cc = a.cc, steps = a.steps;
    const float rcc = rcp_refined(cc);
    const bool fast_cc = a.fast_cc != 0;

    // Position slots as in the binary tile (kernels_fused.h): slots 0 .. KO-1
    // are the tile's own positions (row tid % TY, column tid / TY + k * (NT /
    // TY)), the halo follows (the columns either side, then the rows either
    // side).  Radius-2 flags: bit0 inside, bit1 x>0, bit2 y>0, bit3 i>=1 &&
    // j>=1, bit4 inside the tile (its label is stored here)
    int off2[K2], l2h[K2 - KO];
    const int l2o = (2 + tid / TY) * W2 + 2 + tid % TY;
    uint64_t fl2 = 0;
#pragma unroll
    for (int k = 0; k < K2; ++k) {
        int i, j;
        if (k < KO) {
            i = 2 + tid % TY;
            j = 2 + tid / TY + k * (NT / TY);
        } else {
            const int h = tid + (k - KO) * NT;
            if (h < 4 * TY) {
                const int c = h / TY;
                i = 2 + h % TY;
                j = (c < 2) ? c : H2 - 4 + c;
            } else {
                const int h2 = h - 4 * TY, r4 = h2 & 3;
                i = (r4 < 2) ? r4 : W2 - 4 + r4;
                j = in2(k) ? h2 >> 2 : 0;
            }
            l2h[k - KO] = j * W2 + i;
        }
        const int y = y0 - 2 + i, x = x0 - 2 + j;
        off2[k] = x * ny + y;
        const bool ok = in2(k);
        const bool in = ok && y >= 0 && y < ny && x >= 0 && x < nx;
        const bool own = in && k < KO;
        const uint64_t f = (in ? 1u : 0u) | (x > 0 ? 2u : 0u) | (y > 0 ? 4u : 0u) |
                           ((ok && i >= 1 && j >= 1) ? 8u : 0u) | (own ? 16u : 0u);
        fl2 |= f << (8 * k);
    }
    auto l2 = [&](int k) { return k < KO ? l2o + k * (NT / TY) * W2 : l2h[k - KO]; };
    int off1[K1], l1h[K1 - KO], l2i1h[K1 - KO];
    const int l1o = (1 + tid / TY) * W1 + 1 + tid % TY;
    uint32_t fl1 = 0;
#pragma unroll
    for (int k = 0; k < K1; ++k) {
        int i1, j1;
        if (k < KO) {
            i1 = 1 + tid % TY;
            j1 = 1 + tid / TY + k * (NT / TY);
        } else {
            const int h = tid + (k - KO) * NT;
            if (h < 2 * TY) {
                i1 = 1 + h % TY;
                j1 = (h < TY) ? 0 : H1 - 1;
            } else {
                const int h2 = h - 2 * TY;
                i1 = (h2 & 1) ? W1 - 1 : 0;
                j1 = in1(k) ? h2 >> 1 : 0;
            }
            l1h[k - KO] = j1 * W1 + i1;
            l2i1h[k - KO] = (j1 + 1) * W2 + (i1 + 1);
        }
        const int y = y0 - 1 + i1, x = x0 - 1 + j1;
        off1[k] = x * ny + y;
        const bool in = in1(k) && y >= 0 && y < ny && x >= 0 && x < nx;
        fl1 |= (in ? 1u : 0u) << k;
    }
    auto l1 = [&](int k) { return k < KO ? l1o + k * (NT / TY) * W1 : l1h[k - KO]; };
    auto l2i1 = [&](int k) { return k < KO ? l2(k) : l2i1h[k - KO]; };

    struct Raw {
        float u[K2], pt[K2], pp[K2], bxl[K2], bxr[K2], byl[K2], byr[K2], bzn[K2];
        float al[K1];
    };
    float bz_cur[K2];
#pragma unroll
    for (int k = 0; k < K2; ++k) bz_cur[k] = 0.0f;

    // raw plane q: u, pt, parent pt, own and high x and y faces; the own z face
    // of plane q+1; alpha of plane q-1 (used one step later)
    auto load_raw = [&](int q, Raw &r) {
        const PRef ru = pref(u_in, q, nz, 4, plane);
        const PRef rp = pref(pt, q, nz, 4, plane);
        const PRef rpp = pref(ppt, q, nz, 4, plane);
        const PRef rx = pref(fx, q, nz, 4, plane);
        const PRef ry = pref(fy, q, nz, 4, plane);
#pragma unroll
        for (int k = 0; k < K2; ++k) {
            const int o = off2[k];
            r.u[k] = ldf(ru, o);
            r.pt[k] = ldf(rp, o);
            r.pp[k] = ldf(rpp, o);
            r.bxl[k] = ldf(rx, o);
            r.bxr[k] = ldf(rx, o + ny);
            r.byl[k] = ldf(ry, o);
            r.byr[k] = ldf(ry, o + 1);
        }
        if (IS3D) {
            const PRef rz = pref(fz, q + 1, nz, 4, plane);
#pragma unroll
            for (int k = 0; k < K2; ++k) r.bzn[k] = ldf(rz, off2[k]);
        }
        if (amap) {
            const PRef ra = pref(amap, q - 1, nz, 4, plane);
#pragma unroll
            for (int k = 0; k < K1; ++k) r.al[k] = ldf(ra, off1[k]);
        }
    };

    // S1: labels and g of plane q over the radius-2 region; stepped z face below q
    float gkp[K2];
#pragma unroll
    for (int k = 0; k < K2; ++k) gkp[k] = 0.0f;
    auto s1 = [&](int q, const Raw &r) {
        float *gk = GK[q & 1];
        const int sl = (q + 3) % 3;
        const bool qin = q >= 0 && q < nz;
        const bool qown = q >= z0 && q < z1;
        const PRef wu = pref(u_out, q, nz, 4, plane);
#pragma unroll
        for (int k = 0; k < K2; ++k) {
            const int idx = l2(k);
            const uint32_t f = (uint32_t)(fl2 >> (8 * k));
            const bool live = (f & 1) && qin;
            float d = ((r.byr[k] - r.byl[k]) + r.bxr[k]) - r.bxl[k];
            if (IS3D) d = (d + r.bzn[k]) - bz_cur[k];
            const float un = r.u[k] + cc * ((r.pp[k] - d) - r.pt[k]);
            const float uc = div_fast(un, cc, rcc, fast_cc);
            const float v = live ? steps * (((d + r.pt[k]) - r.pp[k]) - uc) : 0.0f;
            if (k < KO && (f & 16) && qown) st<Io<float>, CP>(wu, off2[k], un);
            if (in2(k)) {
                gk[idx] = v;
                if (IS3D) B1Z[sl][idx] = (live && q > 0) ? (bz_cur[k] + v) - gkp[k] : 0.0f;
            }
            gkp[k] = v;
        }
    };

    // S2a: stepped in-plane faces of plane q from its stored faces fxl, fyl
    auto s2a = [&](int q, const float (&fxl)[K2], const float (&fyl)[K2]) {
        const float *gq = GK[q & 1];
        const int sl = (q + 3) % 3;
        const bool qin = q >= 0 && q < nz;
#pragma unroll
        for (int k = 0; k < K2; ++k) {
            const int idx = l2(k);
            const uint32_t f = (uint32_t)(fl2 >> (8 * k));
            const bool live = (f & 1) && qin;
            if (in2(k)) {
                const float c = gq[idx];
                const float cx = gq[(f & 8) ? idx - W2 : idx];
                const float cy = gq[(f & 8) ? idx - 1 : idx];
                float bx1, by1;
                if (IS3D) {
                    bx1 = (fxl[k] + c) - cx;
                    by1 = (fyl[k] + c) - cy;
                } else {
                    bx1 = fxl[k] + (c - cx);
                    by1 = fyl[k] + (c - cy);
                }
                B1X[sl][idx] = (live && (f & 2)) ? bx1 : 0.0f;
                B1Y[sl][idx] = (live && (f & 4)) ? by1 : 0.0f;
            }
        }
    };

    // S2b: capacity factor of plane q over the radius-1 region, with alpha of plane q
    // in al (a map's values; a scalar alpha otherwise); the own positions' factors
    // also go to g_hi, where the output stage of plane q-1 reads them (same thread)
    float g_lo[KO], g_hi[KO];   // g at the own positions: planes q-1 and q
#pragma unroll
    for (int k = 0; k < KO; ++k) g_lo[k] = g_hi[k] = 1.0f;
    auto s2b = [&](int q, const float (&al_map)[K1]) {
        const int sl = (q + 3) % 3, sn = (q + 4) % 3;
        float *go = G[q & 1];
        const bool qin = q >= 0 && q < nz;
#pragma unroll
        for (int k = 0; k < K1; ++k) {
            const int idx = l1(k);
            const int l = l2i1(k);
            const float xl = B1X[sl][l], xr = B1X[sl][l + W2];
            const float yl = B1Y[sl][l], yr = B1Y[sl][l + 1];
            float t = yl * yl + yr * yr;
            t = (t + xr * xr) + xl * xl;
            if (IS3D) {
                const float zl = B1Z[sl][l], zr = B1Z[sn][l];
                t = (t + zr * zr) + zl * zl;
            }
            const float al = amap ? al_map[k] : as;
            const bool live = ((fl1 >> k) & 1) && qin;
            const float v = live ? hmf_cap(t, al) : 1.0f;
            if (in1(k)) go[idx] = v;
            if (k < KO) g_hi[k] = v;
        }
    };

    float carry[KO];
#pragma unroll
    for (int k = 0; k < KO; ++k) carry[k] = 0.0f;
    auto z_face_up = [&](int s, float gc, float gup, int l2) {
        const int sn = (s + 4) % 3;
        if (!(IS3D && s + 1 < nz)) return 0.0f;
        return (0.5f * B1Z[sn][l2]) * (gup + gc);
    };
    // g of planes s and s+1 at the own positions in g_lo and g_hi
    auto s3_carry = [&](int s) {
#pragma unroll
        for (int k = 0; k < KO; ++k) {
            const int i = tid % TY, j = tid / TY + k * (NT / TY);
            const int l2 = (j + 2) * W2 + (i + 2);
            carry[k] = z_face_up(s, g_lo[k], g_hi[k], l2);
        }
    };
    // S3: rescaled faces and divergence of plane s
    auto s3 = [&](int s) {
        const int sl = (s + 3) % 3;
        const float *g0 = G[s & 1];
        const PRef wd = pref(dvo, s, nz, 4, plane);
        const PRef wx = pref(ox, s, nz, 4, plane);
        const PRef wy = pref(oy, s, nz, 4, plane);
        const PRef wz = pref(IS3D ? oz : ox, s, nz, 4, plane);
#pragma unroll
        for (int k = 0; k < KO; ++k) {
            const int i = tid % TY, j = tid / TY + k * (NT / TY);
            const int y = y0 + i, x = x0 + j;
            const int l2 = (j + 2) * W2 + (i + 2);
            const int l1 = (j + 1) * W1 + (i + 1);
            const float gc = g_lo[k];
            const float bzr = z_face_up(s, gc, g_hi[k], l2);
            float bxl = (0.5f * B1X[sl][l2]) * (gc + g0[l1 - W1]);
            float bxr = (0.5f * B1X[sl][l2 + W2]) * (g0[l1 + W1] + gc);
            float byl = (0.5f * B1Y[sl][l2]) * (gc + g0[l1 - 1]);
            float byr = (0.5f * B1Y[sl][l2 + 1]) * (g0[l1 + 1] + gc);
            bxl = (x > 0) ? bxl : 0.0f;
            bxr = (x + 1 < nx) ? bxr : 0.0f;
            byl = (y > 0) ? byl : 0.0f;
            byr = (y + 1 < ny) ? byr : 0.0f;
            float d = ((byr - byl) + bxr) - bxl;
            float bzl = 0.0f;
            if (IS3D) {
                bzl = carry[k];
                d = (d + bzr) - bzl;
                carry[k] = bzr;
            }
            if (y < ny && x < nx) {
                const int off = x * ny + y;
                st<Io<float>, CP>(wd, off, d);
                st<Io<float>, CP>(wx, off, bxl);
                st<Io<float>, CP>(wy, off, byl);
                if (IS3D) st<Io<float>, CP>(wz, off, bzl);
            }
        }
    };

    // One barrier per plane step.  Step t: the loads of raw plane t+2, S2b(t), S3(t-1),
    // S2a(t+1), S1(t+2); barrier.  Each stage reads in LDS only what an earlier step
    // wrote, and each ring slot it writes was last read in an earlier step.  In
    // registers instead: g of plane t at S3(t-1)'s own positions (computed by the same
    // thread in S2b(t)), alpha of plane t for S2b and the stored bx, by of plane t+1 for
    // S2a, taken from the raw plane S1 read one step earlier (alc, fxl, fyl).
    float alc[K1], fxl[K2], fyl[K2];
    auto keep_raw = [&](const Raw &rr) {
#pragma unroll
        for (int k = 0; k < K1; ++k) alc[k] = rr.al[k];
#pragma unroll
        for (int k = 0; k < K2; ++k) {
            fxl[k] = rr.bxl[k];
            fyl[k] = rr.byl[k];
        }
    };
    auto step = [&](int t, Raw &r) {
        if (t + 2 <= z1 + 1) load_raw(t + 2, r);
        if (t >= z0 - 1) {
#pragma unroll
            for (int k = 0; k < KO; ++k) g_lo[k] = g_hi[k];
            s2b(t, alc);
        }
        if (t - 1 >= z0) s3(t - 1);
        else if (t - 1 == z0 - 1) s3_carry(t - 1);
        if (t >= z0 - 2 && t + 1 <= z1) s2a(t + 1, fxl, fyl);
        if (t + 2 <= z1 + 1) {
            s1(t + 2, r);
            keep_raw(r);
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < K2; ++k) bz_cur[k] = r.bzn[k];
    };
    Raw r;
#pragma unroll
    for (int k = 0; k < K2; ++k) r.bzn[k] = 0.0f;
#pragma unroll
    for (int k = 0; k < K1; ++k) r.al[k] = alc[k] = as;
    if (IS3D) {
        const PRef rz = pref(fz, z0 - 2, nz, 4, plane);
#pragma unroll
        for (int k = 0; k < K2; ++k) bz_cur[k] = ldf(rz, off2[k]);
    }
    if constexpr (IS3D) {
        for (int t = z0 - 4; t <= z1; ++t) step(t, r);
    } else {
        // one plane: its stages in dependency order, no z pipeline
        load_raw(0, r);
        if (amap) {
            const PRef ra = pref(amap, 0, nz, 4, plane);
#pragma unroll
            for (int k = 0; k < K1; ++k) r.al[k] = ldf(ra, off1[k]);
        }
        s1(0, r);
        __syncthreads();
        s2a(0, r.bxl, r.byl);
        __syncthreads();
        s2b(0, r.al);
#pragma unroll
        for (int k = 0; k < KO; ++k) g_lo[k] = g_hi[k];   // S3(0) reads g of plane 0
        __syncthreads();
        s3(0);
    }
}

template <bool IS3D, int TX, int NT, int WAVES>
__global__ __launch_bounds__(NT, WAVES) void kf_hmf_flow(HmfArgs a) {
    const int b = xcd_remap((int)blockIdx.x, (int)gridDim.x);
    if (b >= a.nblocks) return;
    hmf_tile<IS3D, TX, NT>(a, b);
}

// ---- per-pass kernels -------------------------------------------------------
// Grid: (cdiv(N, 256), nflow); blockIdx.y is the slot of a non-root node.  g holds
// one field per flow slot (the gradient step's g, later the capacity factors);
// shape-constrained nodes own slots of a (exemption amounts), ex (per face) and
// the normalised directions L.
struct HmfPass {
    const HmfNode *nodes;
    const int32_t *flow;
    float *u, *pt, *div;     // [M][N]
    float *f[3];             // [M][N] each
    float *g;                // [nflow][N]
    float *a;                // [nshape][N]
    float *ex;               // [nshape][3][N]
    const float *L;          // [nshape][3][N]
    int64_t N;
    int ny, nx, nz;
    float cc, steps;
    int fast_cc;
};

struct HmfAt {
    int n, s;        // node, flow slot
    int64_t g, on;   // voxel, node offset
    int c[3];        // coordinates (y, x, z)
};

__device__ __forceinline__ bool hmf_at(const HmfPass &p, HmfAt &w) {
    w.g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (w.g >= p.N) return false;
    w.s = blockIdx.y;
    w.n = p.flow[w.s];
    w.on = (int64_t)w.n * p.N;
    const int64_t pl = (int64_t)p.ny * p.nx;
    w.c[0] = (int)(w.g % p.ny);
    w.c[1] = (int)((w.g / p.ny) % p.nx);
    w.c[2] = (int)(w.g / pl);
    return true;
}

// labels, then g (asetsHMF3D.m:193, :209)
__global__ void kh_labels_g(HmfPass p) {
    HmfAt w;
    if (!hmf_at(p, w)) return;
    const int64_t i = w.on + w.g;
    const float pp = p.pt[(int64_t)p.nodes[w.n].parent * p.N + w.g];
    const float d = p.div[i], t = p.pt[i];
    const float un = p.u[i] + p.cc * ((pp - d) - t);
    p.u[i] = un;
    const float uc = div_fast(un, p.cc, rcp_refined(p.cc), p.fast_cc != 0);
    p.g[(int64_t)w.s * p.N + w.g] = p.steps * (((d + t) - pp) - uc);
}

template <bool IS3D>
__global__ void kh_step(HmfPass p) {
    HmfAt w;
    if (!hmf_at(p, w)) return;
    const float *g = p.g + (int64_t)w.s * p.N;
    const int64_t st[3] = {1, p.ny, (int64_t)p.ny * p.nx};
    const float gc = g[w.g];
#pragma unroll
    for (int k = 0; k < (IS3D ? 3 : 2); ++k) {
        if (w.c[k] == 0) continue;
        float *f = p.f[k] + w.on + w.g;
        *f = IS3D ? (*f + gc) - g[w.g - st[k]] : *f + (gc - g[w.g - st[k]]);
    }
}

// own face k of voxel g and the high face (the next voxel's own face), 0 outside
template <bool IS3D>
__device__ __forceinline__ void hmf_faces(const HmfPass &p, const HmfAt &w, int k, float &lo,
                                          float &hi) {
    const int64_t st[3] = {1, p.ny, (int64_t)p.ny * p.nx};
    const int ext[3] = {p.ny, p.nx, p.nz};
    const float *f = p.f[k] + w.on;
    lo = (w.c[k] > 0) ? f[w.g] : 0.0f;
    hi = (w.c[k] + 1 < ext[k]) ? f[w.g + st[k]] : 0.0f;
}

// exemption amount per voxel (asetsHMF3D.m:216-223)
template <bool IS3D>
__global__ void kh_exempt_a(HmfPass p) {
    HmfAt w;
    if (!hmf_at(p, w)) return;
    const int sh = p.nodes[w.n].shape;
    if (sh < 0) return;
    const float *L = p.L + (int64_t)sh * 3 * p.N;
    float a = 0.0f;
#pragma unroll
    for (int k = 0; k < (IS3D ? 3 : 2); ++k) {
        float lo, hi, t;
        hmf_faces<IS3D>(p, w, k, lo, hi);
        const float lk = L[(int64_t)k * p.N + w.g];
        if (k == 0) {
            t = (lo < 0.0f) ? lo * lk : 0.0f;
            a = (t > 0.0f) ? t : 0.0f;
            t = (hi > 0.0f) ? hi * lk : 0.0f;
            a = a + ((t > 0.0f) ? t : 0.0f);
        } else {
            t = (hi > 0.0f) ? hi * lk : 0.0f;
            a = a + ((t > 0.0f) ? t : 0.0f);
            t = (lo < 0.0f) ? lo * lk : 0.0f;
            a = a + ((t > 0.0f) ? t : 0.0f);
        }
    }
    p.a[(int64_t)sh * p.N + w.g] = a;
}

// exemption per face and its removal (:226-238)
template <bool IS3D>
__global__ void kh_exempt_apply(HmfPass p) {
    HmfAt w;
    if (!hmf_at(p, w)) return;
    const int sh = p.nodes[w.n].shape;
    if (sh < 0) return;
    const int64_t st[3] = {1, p.ny, (int64_t)p.ny * p.nx};
    const float *L = p.L + (int64_t)sh * 3 * p.N;
    const float *A = p.a + (int64_t)sh * p.N;
    float *EX = p.ex + (int64_t)sh * 3 * p.N;
#pragma unroll
    for (int k = 0; k < (IS3D ? 3 : 2); ++k) {
        if (w.c[k] == 0) continue;
        const int64_t lo = w.g - st[k];
        const float *Lk = L + (int64_t)k * p.N;
        float *f = p.f[k] + w.on + w.g;
        const float q = *f;
        float e = 0.0f;
        if (q > 0.0f && Lk[lo] > 0.0f) e = Lk[lo] * A[lo];
        else if (q < 0.0f && Lk[w.g] < 0.0f) e = Lk[w.g] * A[w.g];
        EX[(int64_t)k * p.N + w.g] = e;
        *f = q - e;
    }
}

// capacity factor per voxel (:240-258)
template <bool IS3D>
__global__ void kh_proj(HmfPass p) {
    HmfAt w;
    if (!hmf_at(p, w)) return;
    float t = 0.0f;
#pragma unroll
    for (int k = 0; k < (IS3D ? 3 : 2); ++k) {
        float lo, hi;
        hmf_faces<IS3D>(p, w, k, lo, hi);
        if (k == 0) t = lo * lo + hi * hi;
        else t = (t + hi * hi) + lo * lo;
    }
    const HmfNode nd = p.nodes[w.n];
    const float al = nd.alpha ? nd.alpha[w.g] : nd.alpha_s;
    p.g[(int64_t)w.s * p.N + w.g] = hmf_cap(t, al);
}

// rescaled faces (:259-261 / :283-285)
template <bool IS3D>
__global__ void kh_rescale(HmfPass p) {
    HmfAt w;
    if (!hmf_at(p, w)) return;
    const int64_t st[3] = {1, p.ny, (int64_t)p.ny * p.nx};
    const float *g = p.g + (int64_t)w.s * p.N;
    const int sh = p.nodes[w.n].shape;
    const float gc = g[w.g];
#pragma unroll
    for (int k = 0; k < (IS3D ? 3 : 2); ++k) {
        if (w.c[k] == 0) continue;
        float *f = p.f[k] + w.on + w.g;
        const float r = (0.5f * *f) * (gc + g[w.g - st[k]]);
        *f = (sh >= 0) ? p.ex[((int64_t)sh * 3 + k) * p.N + w.g] + r : r;
    }
}

template <bool IS3D>
__global__ void kh_div(HmfPass p) {
    HmfAt w;
    if (!hmf_at(p, w)) return;
    float d = 0.0f;
#pragma unroll
    for (int k = 0; k < (IS3D ? 3 : 2); ++k) {
        float lo, hi;
        hmf_faces<IS3D>(p, w, k, lo, hi);
        d = (k == 0) ? hi - lo : (d + hi) - lo;
    }
    p.div[w.on + w.g] = d;
}

// Shape directions normalised where |l| > 0.001, else 0 (InitializeFullFlow,
// asetsHMF3D.m:117-128).  src: per slot the ndim caller fields; dst [slot][3][N].
struct HmfShapeSrc {
    const float *l[3];
};
template <bool IS3D>
__global__ void k_hmf_lnorm(const HmfShapeSrc *src, float *dst, int64_t N) {
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= N) return;
    const HmfShapeSrc s = src[blockIdx.y];
    float q[3] = {0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int k = 0; k < (IS3D ? 3 : 2); ++k) q[k] = s.l[k][g];
    float s2 = q[0] * q[0] + q[1] * q[1];
    if (IS3D) s2 = s2 + q[2] * q[2];
    const float den = __builtin_sqrtf(s2);
    float *o = dst + (int64_t)blockIdx.y * 3 * N;
#pragma unroll
    for (int k = 0; k < 3; ++k) o[(int64_t)k * N + g] = (den > 0.001f) ? q[k] / den : 0.0f;
}

// ---- costs (CalculateCostsFullFlow, asetsHMF2D.m:184-203) -----------------------
// Per node and block: (sum of u*Ct over a leaf, sum of u*div below the root), the
// fp32 products summed in FP64; then one workgroup per node adds its blocks in order.
constexpr int kHmfCostBlocks = 256;
__global__ __launch_bounds__(256) void k_hmf_cost_part(const HmfNode *nodes, const float *u,
                                                       const float *div, int64_t N,
                                                       double *part) {
    __shared__ double red[16];
    const int n = blockIdx.y;
    const HmfNode nd = nodes[n];
    const float *un = u + (int64_t)n * N, *dn = div + (int64_t)n * N;
    double dc = 0.0, rc = 0.0;
    for (int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; g < N;
         g += (int64_t)gridDim.x * blockDim.x) {
        const float ug = un[g];
        if (nd.Ct) dc += (double)(ug * nd.Ct[g]);
        if (nd.parent >= 0) rc += (double)(ug * dn[g]);
    }
    const double sd = block_sum(dc, red);
    __syncthreads();
    const double sr = block_sum(rc, red);
    if (threadIdx.x == 0) {
        part[((int64_t)n * gridDim.x + blockIdx.x) * 2] = sd;
        part[((int64_t)n * gridDim.x + blockIdx.x) * 2 + 1] = sr;
    }
}

__global__ __launch_bounds__(256) void k_hmf_cost_final(const double *part, int nb,
                                                        double *out) {
    __shared__ double red[16];
    const int n = blockIdx.x;
    double d = 0.0, r = 0.0;
    for (int b = threadIdx.x; b < nb; b += blockDim.x) {
        d += part[((int64_t)n * nb + b) * 2];
        r += part[((int64_t)n * nb + b) * 2 + 1];
    }
    const double sd = block_sum(d, red);
    __syncthreads();
    const double sr = block_sum(r, red);
    if (threadIdx.x == 0) {
        out[2 * n] = sd;
        out[2 * n + 1] = sr;
    }
}

}  // namespace asets
