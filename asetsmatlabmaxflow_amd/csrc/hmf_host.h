// hmf_host.h -- host side of the hierarchical max-flow solve (asets_hmf_dev).
// Included by asets_hip.hip inside its anonymous namespace: uses its workspace
// cache, error reporting and fill helpers.
//
// MaxFullFlow (asetsHMF3D.m:77-87): InitializeFullFlow (zero state, shape
// directions normalised), numIts iterations of k_hmf_sink + the spatial step
// (kf_hmf_flow, or the kh_* passes), then CalculateCostsFullFlow.

// One launch for a solve's output copies: segment blockIdx.y copies n floats from src
// to dst (src null: fills 0).
struct CopySeg {
    float *dst;
    const float *src;
    int64_t n;
};
constexpr int kMaxCopySegsPerLaunch = 65535;   // grid height

__global__ void k_multi_copy(const CopySeg *segs) {
    const CopySeg sg = segs[blockIdx.y];
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < sg.n; i += stride)
        sg.dst[i] = sg.src ? sg.src[i] : 0.0f;
}

// Enqueue the segments through the device table d_segs (room for segs.size()).
inline hipError_t multi_copy(const std::vector<CopySeg> &segs, CopySeg *d_segs, hipStream_t s) {
    if (segs.empty()) return hipSuccess;
    const hipError_t e = hipMemcpyAsync(d_segs, segs.data(), segs.size() * sizeof(CopySeg),
                                        hipMemcpyHostToDevice, s);
    if (e != hipSuccess) return e;
    int64_t nmax = 0;
    for (const CopySeg &c : segs) nmax = std::max(nmax, c.n);
    const unsigned gx = (unsigned)std::max<int64_t>(std::min<int64_t>(cdiv(nmax, kBlock), 1024), 1);
    for (size_t i0 = 0; i0 < segs.size(); i0 += kMaxCopySegsPerLaunch) {
        const unsigned gy = (unsigned)std::min<size_t>(segs.size() - i0, kMaxCopySegsPerLaunch);
        k_multi_copy<<<dim3(gx, gy), kBlock, 0, s>>>(d_segs + i0);
    }
    return hipGetLastError();
}

struct HmfPlan {
    std::vector<HmfNode> nodes;
    std::vector<int32_t> flow;      // non-root nodes in index order
    std::vector<HmfEvent> events;   // depth-first sink program
    std::vector<int32_t> bfs;       // breadth-first order from the root
    std::vector<std::vector<int32_t>> kids;
    std::vector<HmfShapeSrc> shape;
    int root = -1;
    int levels = 0;                 // internal-node depths + 1 (sink stack)
};

constexpr int kHmfMaxNodes = 4096;
constexpr int kHmfMaxLevels = 16;

asets_status hmf_plan(int ndim, const asets_hmf_node *in, int32_t M, HmfPlan &pl) {
    pl.kids.assign(M, {});
    pl.nodes.resize(M);
    for (int32_t n = 0; n < M; ++n) {
        const int32_t p = in[n].parent;
        if (p < 0) {
            if (pl.root >= 0) return fail(ASETS_EBADARG, "HMF tree has more than one root");
            pl.root = n;
        } else {
            if (p >= M || p == n)
                return fail(ASETS_EBADARG, "node %d: bad parent index %d", n, p);
            pl.kids[p].push_back(n);
        }
    }
    if (pl.root < 0) return fail(ASETS_EBADARG, "HMF tree has no root");
    if (pl.kids[pl.root].empty()) return fail(ASETS_EBADARG, "the HMF root has no children");
    std::vector<int32_t> depth(M, -1);
    depth[pl.root] = 0;
    pl.bfs.push_back(pl.root);
    for (size_t k = 0; k < pl.bfs.size(); ++k)
        for (int32_t c : pl.kids[pl.bfs[k]]) {
            if (depth[c] >= 0) return fail(ASETS_EBADARG, "HMF tree has a cycle");
            depth[c] = depth[pl.bfs[k]] + 1;
            pl.bfs.push_back(c);
        }
    if ((int32_t)pl.bfs.size() != M)
        return fail(ASETS_EBADARG, "HMF tree: %d nodes are not reachable from the root",
                    M - (int32_t)pl.bfs.size());
    for (int32_t n = 0; n < M; ++n) {
        const bool leaf = pl.kids[n].empty();
        if (!leaf && depth[n] + 1 > kHmfMaxLevels)
            return fail(ASETS_ETOOBIG, "HMF tree deeper than %d levels of internal nodes",
                        kHmfMaxLevels);
        if (!leaf) pl.levels = std::max(pl.levels, depth[n] + 1);
        if (leaf && !in[n].Ct) return fail(ASETS_EBADARG, "leaf node %d has no Ct", n);
        HmfNode &h = pl.nodes[n];
        h.Ct = leaf ? in[n].Ct : nullptr;
        h.alpha = in[n].alpha_map;
        h.alpha_s = in[n].alpha;
        h.parent = in[n].parent;
        h.shape = -1;
        h.pad = 0;
        if (n == pl.root) {
            h.alpha = nullptr;   // the root has no spatial flows
            continue;
        }
        pl.flow.push_back(n);
        int nl = 0;
        for (int k = 0; k < ndim; ++k) nl += in[n].l[k] ? 1 : 0;
        if (nl != 0 && nl != ndim)
            return fail(ASETS_EBADARG, "node %d: give all %d shape directions or none", n, ndim);
        if (nl == ndim) {
            h.shape = (int32_t)pl.shape.size();
            HmfShapeSrc s{};
            for (int k = 0; k < ndim; ++k) s.l[k] = in[n].l[k];
            pl.shape.push_back(s);
        }
    }
    // depth-first program, children in index order
    std::function<void(int32_t)> visit = [&](int32_t n) {
        if (pl.kids[n].empty()) {
            pl.events.push_back({kHmfLeaf, n, depth[n], 0.0f, 1});
            return;
        }
        pl.events.push_back({kHmfEnter, n, depth[n], 0.0f, 0});
        for (int32_t c : pl.kids[n]) visit(c);
        const float cnt = (float)(pl.kids[n].size() + (n == pl.root ? 0 : 1));
        pl.events.push_back({kHmfExit, n, depth[n], cnt, 0});
    };
    visit(pl.root);
    // runs of consecutive sibling leaves (the kernel loads four at a time)
    for (int e = (int)pl.events.size() - 2; e >= 0; --e) {
        HmfEvent &a = pl.events[e];
        const HmfEvent &b = pl.events[e + 1];
        if (a.type == kHmfLeaf && b.type == kHmfLeaf && in[a.node].parent == in[b.node].parent)
            a.run = b.run + 1;
    }
    return ASETS_OK;
}

// z-chunk length of the fused flow kernel with 64 x 8 or 64 x 32 tiles (the
// 64 x 16 tile uses pick_zchunk): a workgroup streams chunk + 4 planes
// (the 4-plane prologue), and the grid runs in rounds of `slots` resident
// workgroups, so the time is about rounds(tiles * chunks) * (chunk + 4).  The
// smallest such cost wins; ties go to the longer chunk (fewer halo re-reads).
int hmf_zchunk(int64_t tiles, int64_t nz, int64_t slots) {
    static const int cand[] = {8, 12, 16, 24, 32, 48, 64, 96, 128, 192, 256, 384, 512};
    int64_t best = -1, best_zc = nz;
    auto consider = [&](int64_t zc) {
        const int64_t nb = tiles * cdiv(nz, zc);
        const int64_t cost = cdiv(nb, slots) * (zc + 4);
        if (best < 0 || cost < best || (cost == best && zc > best_zc)) {
            best = cost;
            best_zc = zc;
        }
    };
    for (int zc : cand)
        if (zc < nz) consider(zc);
    consider(nz);
    return (int)best_zc;
}

template <bool IS3D, int TX, int NT>
int hmf_resident_t() {
    constexpr int W = (NT >= 512) ? 4 : 3;
    int per_cu = 0, cus = 0, dev = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kf_hmf_flow<IS3D, TX, NT, W>, NT,
                                                     0) != hipSuccess ||
        hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) {
        (void)hipGetLastError();
        return 512;
    }
    return std::max(1, per_cu) * std::max(1, cus);
}

int hmf_resident(bool is3d, int tx) {
    if (is3d) {
        if (tx == 8) return hmf_resident_t<true, 8, 256>();
        if (tx == 32) return hmf_resident_t<true, 32, 1024>();
        return hmf_resident_t<true, 16, 512>();
    }
    if (tx == 8) return hmf_resident_t<false, 8, 256>();
    if (tx == 32) return hmf_resident_t<false, 32, 1024>();
    return hmf_resident_t<false, 16, 512>();
}

template <bool IS3D, int TX, int NT>
void launch_hmf_flow_t(const HmfArgs &a, int grid, hipStream_t s) {
    constexpr int W = (NT >= 512) ? 4 : 3;
    kf_hmf_flow<IS3D, TX, NT, W><<<grid, NT, 0, s>>>(a);
}

void launch_hmf_flow(bool is3d, int tx, const HmfArgs &a, int grid, hipStream_t s) {
    if (is3d) {
        if (tx == 8) launch_hmf_flow_t<true, 8, 256>(a, grid, s);
        else if (tx == 32) launch_hmf_flow_t<true, 32, 1024>(a, grid, s);
        else launch_hmf_flow_t<true, 16, 512>(a, grid, s);
    } else {
        if (tx == 8) launch_hmf_flow_t<false, 8, 256>(a, grid, s);
        else if (tx == 32) launch_hmf_flow_t<false, 32, 1024>(a, grid, s);
        else launch_hmf_flow_t<false, 16, 512>(a, grid, s);
    }
}

template <bool IS3D>
void launch_hmf_passes(const HmfPass &p, int nflow, bool shape, hipStream_t s) {
    const dim3 grid((unsigned)cdiv(p.N, kBlock), (unsigned)nflow);
    kh_labels_g<<<grid, kBlock, 0, s>>>(p);
    kh_step<IS3D><<<grid, kBlock, 0, s>>>(p);
    if (shape) {
        kh_exempt_a<IS3D><<<grid, kBlock, 0, s>>>(p);
        kh_exempt_apply<IS3D><<<grid, kBlock, 0, s>>>(p);
    }
    kh_proj<IS3D><<<grid, kBlock, 0, s>>>(p);
    kh_rescale<IS3D><<<grid, kBlock, 0, s>>>(p);
    kh_div<IS3D><<<grid, kBlock, 0, s>>>(p);
}

asets_status hmf_dev(int ndim, int64_t ny, int64_t nx, int64_t nz, asets_hmf_node *in,
                     int32_t M, int32_t num_iter, float steps, double cc_d, double *seconds,
                     const asets_opts *o) {
    auto t_start = std::chrono::steady_clock::now();
    // MATLAB's cc is a double: fp32 operations use (float)cc; the root's 1/cc .*
    // ones(h.D) (asetsHMF3D.m:320) is a double array rounded where it meets fp32 data
    const float cc = (float)cc_d;
    if (ndim != 2 && ndim != 3) return fail(ASETS_EBADARG, "ndim must be 2 or 3");
    if (ndim == 2 && nz != 1) return fail(ASETS_EBADARG, "2D solve needs nz == 1");
    if (!in) return fail(ASETS_EBADARG, "nodes is NULL");
    if (M < 2 || M > kHmfMaxNodes)
        return fail(ASETS_EBADARG, "an HMF tree has 2 to %d nodes (got %d)", kHmfMaxNodes, M);
    if (num_iter < 0) return fail(ASETS_EBADARG, "numIts must be >= 0 (got %d)", num_iter);
    if (!(cc > 0.0f) && !(cc < 0.0f)) return fail(ASETS_EBADARG, "cc must be non-zero");
    if (o && o->flow_dtype != ASETS_FLOW_F32)
        return fail(ASETS_EBADARG, "fp16 flow storage is available for the binary model only");
    asets_status st = check_dims(ny, nx, nz, M);
    if (st) return st;
    HmfPlan pl;
    st = hmf_plan(ndim, in, M, pl);
    if (st) return st;
    const bool is3d = (ndim == 3);
    const int nd = ndim;
    const int nflow = (int)pl.flow.size();
    const int nshape = (int)pl.shape.size();
    const int kern = o ? o->kernel : ASETS_KERNEL_AUTO;
    bool fused = (kern != ASETS_KERNEL_PASSES) && nshape == 0 && fused_fits(ny, nx, nz);
    if (kern == ASETS_KERNEL_FUSED && !fused) {
        if (nshape)
            return fail(ASETS_EBADARG, "nodes with a shape constraint run the per-pass kernels "
                                       "only (the exemption step has dependency radius 3)");
        return fail(ASETS_ETOOBIG, "volume too large for the fused kernel (fields of 2^31 "
                                   "bytes or more); use the per-pass kernels");
    }

    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return fail(ASETS_EDEVICE, "device id %d", dev);
    DeviceCache &dc = g_cache[dev];
    std::lock_guard<std::mutex> lock(dc.mu);
    hipStream_t s = o ? (hipStream_t)o->stream : nullptr;

    const int64_t N = ny * nx * nz, MN = (int64_t)M * N;
    const int nbuf = fused ? 2 : 1;
    const int nev = (int)pl.events.size();
    size_t need = Carver::sz(M, sizeof(HmfNode)) + Carver::sz(nflow, 4) +
                  Carver::sz(nev, sizeof(HmfEvent)) + Carver::sz(std::max(nshape, 1), sizeof(HmfShapeSrc)) +
                  Carver::sz((int64_t)M * kHmfCostBlocks * 2, 8) + Carver::sz(2 * M, 8) +
                  (size_t)(nbuf * (1 + nd) + 2) * Carver::sz(MN, 4) +
                  Carver::sz(std::max<int64_t>((int64_t)M * 9, 16), sizeof(CopySeg));
    if (!fused) need += Carver::sz((int64_t)nflow * N, 4);
    if (nshape) need += Carver::sz((int64_t)nshape * N, 4) + 2 * Carver::sz((int64_t)nshape * 3 * N, 4);
    st = ensure(&dc.dev, &dc.dev_bytes, need);
    if (st) return st;
    Carver cv(dc.dev);
    HmfNode *d_nodes = cv.take<HmfNode>(M);
    int32_t *d_flow = cv.take<int32_t>(nflow);
    HmfEvent *d_ev = cv.take<HmfEvent>(nev);
    HmfShapeSrc *d_shape = cv.take<HmfShapeSrc>(std::max(nshape, 1));
    double *d_cpart = cv.take<double>((int64_t)M * kHmfCostBlocks * 2);
    double *d_cost = cv.take<double>(2 * M);
    // segment tables of the zero fills at the start (up to 11) and of the output copies
    // (u, pt, div, 3 faces, 3 shape directions per node)
    CopySeg *d_segs = cv.take<CopySeg>(std::max<int64_t>((int64_t)M * 9, 16));
    float *ub[2] = {nullptr, nullptr}, *fb[2][3] = {};
    for (int b = 0; b < nbuf; ++b) {
        ub[b] = cv.take<float>(MN);
        for (int k = 0; k < nd; ++k) fb[b][k] = cv.take<float>(MN);
    }
    float *pt = cv.take<float>(MN);
    float *dv = cv.take<float>(MN);
    float *gbuf = fused ? nullptr : cv.take<float>((int64_t)nflow * N);
    float *abuf = nullptr, *exbuf = nullptr, *Lbuf = nullptr;
    if (nshape) {
        abuf = cv.take<float>((int64_t)nshape * N);
        exbuf = cv.take<float>((int64_t)nshape * 3 * N);
        Lbuf = cv.take<float>((int64_t)nshape * 3 * N);
    }

    HIPCHK(hipMemcpyAsync(d_nodes, pl.nodes.data(), M * sizeof(HmfNode), hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(d_flow, pl.flow.data(), nflow * sizeof(int32_t), hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(d_ev, pl.events.data(), nev * sizeof(HmfEvent), hipMemcpyHostToDevice, s));
    if (nshape)
        HIPCHK(hipMemcpyAsync(d_shape, pl.shape.data(), nshape * sizeof(HmfShapeSrc),
                              hipMemcpyHostToDevice, s));
    // InitializeFullFlow: zero state (asetsHMF3D.m:109-115) in one launch, normalised
    // directions
    {
        std::vector<CopySeg> zero;
        for (int b = 0; b < nbuf; ++b) {
            zero.push_back({ub[b], nullptr, MN});
            for (int k = 0; k < nd; ++k) zero.push_back({fb[b][k], nullptr, MN});
        }
        zero.push_back({pt, nullptr, MN});
        zero.push_back({dv, nullptr, MN});
        if (nshape) zero.push_back({exbuf, nullptr, (int64_t)nshape * 3 * N});
        HIPCHK(multi_copy(zero, d_segs, s));
    }
    if (nshape) {
        const dim3 grid((unsigned)cdiv(N, kBlock), (unsigned)nshape);
        if (is3d) k_hmf_lnorm<true><<<grid, kBlock, 0, s>>>(d_shape, Lbuf, N);
        else k_hmf_lnorm<false><<<grid, kBlock, 0, s>>>(d_shape, Lbuf, N);
    }
    HIPCHK(hipGetLastError());

    const int fast_cc = (std::fabs(cc) >= 0x1p-16f && std::fabs(cc) <= 0x1p16f) ? 1 : 0;
    const float one_cc = (float)(1.0 / cc_d);
    const size_t sink_lds = (size_t)pl.levels * 4 * kBlock * sizeof(float);
    const unsigned nbv = (unsigned)cdiv(N, kBlock);

    HmfArgs fa{};
    HmfPass pa{};
    int grid = 0, tx = 16;
    if (fused) {
        tx = (o && o->tile_x > 0) ? o->tile_x : 16;
        if (tx != 8 && tx != 32) tx = 16;
        const int64_t tiles = cdiv(ny, 64) * cdiv(nx, tx) * nflow;
        // the default 64 x 16 tile: the binary kernel's per-CU load model (192^3
        // two-level tree: 64 -> 32-plane chunks, 0.967 -> 0.910 ms per iteration)
        const int zc = (o && o->zchunk > 0) ? (int)std::min<int64_t>(o->zchunk, nz)
                       : (tx == 16) ? pick_zchunk(tiles, nz)
                                    : hmf_zchunk(tiles, nz, hmf_resident(is3d, tx));
        fa.nodes = d_nodes;
        fa.flow = d_flow;
        fa.nflow = nflow;
        fa.pt = pt;
        fa.div = dv;
        fa.N = N;
        fa.ny = (int)ny;
        fa.nx = (int)nx;
        fa.nz = (int)nz;
        fa.plane = (int)(ny * nx);
        fa.cc = cc;
        fa.steps = steps;
        fa.fast_cc = fast_cc;
        fa.zchunk = zc;
        fa.ntile_y = (int)cdiv(ny, 64);
        fa.ntile_x = (int)cdiv(nx, tx);
        fa.nzc = (int)cdiv(nz, zc);
        const int64_t nb = (int64_t)fa.ntile_y * fa.ntile_x * fa.nzc * nflow;
        if (nb > INT32_MAX / 2) return fail(ASETS_ETOOBIG, "too many work items");
        fa.nblocks = (int)nb;
        grid = (int)(cdiv(nb, 8) * 8);
    } else {
        pa.nodes = d_nodes;
        pa.flow = d_flow;
        pa.u = ub[0];
        pa.pt = pt;
        pa.div = dv;
        for (int k = 0; k < 3; ++k) pa.f[k] = k < nd ? fb[0][k] : nullptr;
        pa.g = gbuf;
        pa.a = abuf;
        pa.ex = exbuf;
        pa.L = Lbuf;
        pa.N = N;
        pa.ny = (int)ny;
        pa.nx = (int)nx;
        pa.nz = (int)nz;
        pa.cc = cc;
        pa.steps = steps;
        pa.fast_cc = fast_cc;
    }

    EventPair evp;
    hipEvent_t &ev0 = evp.a, &ev1 = evp.b;
    if (o && o->kernel_ms) {
        HIPCHK(hipEventCreate(&ev0));
        HIPCHK(hipEventCreate(&ev1));
        HIPCHK(hipEventRecord(ev0, s));
    }
    for (int32_t it = 0; it < num_iter; ++it) {
        const int cur = fused ? (it & 1) : 0;
        k_hmf_sink<<<nbv, kBlock, sink_lds, s>>>(d_nodes, d_ev, nev, ub[cur], pt, dv, N, cc,
                                                 fast_cc, one_cc);
        if (fused) {
            fa.u_in = ub[cur];
            fa.u_out = ub[cur ^ 1];
            for (int k = 0; k < 3; ++k) {
                fa.f_in[k] = k < nd ? fb[cur][k] : nullptr;
                fa.f_out[k] = k < nd ? fb[cur ^ 1][k] : nullptr;
            }
            launch_hmf_flow(is3d, tx, fa, grid, s);
        } else if (is3d) {
            launch_hmf_passes<true>(pa, nflow, nshape > 0, s);
        } else {
            launch_hmf_passes<false>(pa, nflow, nshape > 0, s);
        }
        if ((it & 63) == 63) HIPCHK(hipGetLastError());
    }
    HIPCHK(hipGetLastError());
    if (ev1) HIPCHK(hipEventRecord(ev1, s));
    const int fin = fused ? (num_iter & 1) : 0;

    // CalculateCostsFullFlow: own sums per node, subtree totals on the host
    k_hmf_cost_part<<<dim3(kHmfCostBlocks, (unsigned)M), kBlock, 0, s>>>(d_nodes, ub[fin], dv, N,
                                                                       d_cpart);
    k_hmf_cost_final<<<(unsigned)M, kBlock, 0, s>>>(d_cpart, kHmfCostBlocks, d_cost);
    HIPCHK(hipGetLastError());
    std::vector<double> cost(2 * M);
    HIPCHK(hipMemcpyAsync(cost.data(), d_cost, 2 * M * sizeof(double), hipMemcpyDeviceToHost, s));
    // outputs, every copy and fill in one launch (a 40-leaf tree has about 200)
    std::vector<CopySeg> segs;
    segs.reserve((size_t)M * 9);
    for (int32_t n = 0; n < M; ++n) {
        const int64_t on = (int64_t)n * N;
        if (in[n].u) segs.push_back({in[n].u, ub[fin] + on, N});
        if (in[n].pt) segs.push_back({in[n].pt, pt + on, N});
        if (in[n].div) segs.push_back({in[n].div, dv + on, N});
        for (int k = 0; k < 3; ++k)
            if (in[n].faces[k]) segs.push_back({in[n].faces[k], k < nd ? fb[fin][k] + on : nullptr, N});
        const int sh = pl.nodes[n].shape;
        for (int k = 0; k < 3; ++k)
            if (in[n].l_out[k])
                segs.push_back({in[n].l_out[k],
                                sh >= 0 ? Lbuf + ((int64_t)sh * 3 + k) * N : nullptr, N});
    }
    HIPCHK(multi_copy(segs, d_segs, s));
    HIPCHK(hipStreamSynchronize(s));
    for (int32_t n = 0; n < M; ++n) {
        in[n].data_cost = cost[2 * n];
        in[n].regu_cost = cost[2 * n + 1];
    }
    for (int k = M - 1; k >= 0; --k) {
        const int32_t n = pl.bfs[k];
        for (int32_t c : pl.kids[n]) {
            in[n].data_cost += in[c].data_cost;
            in[n].regu_cost += in[c].regu_cost;
        }
    }
    if (ev0) {
        float ms = 0.0f;
        HIPCHK(hipEventElapsedTime(&ms, ev0, ev1));
        *o->kernel_ms = num_iter > 0 ? (double)ms / num_iter : 0.0;
    }
    if (seconds)
        *seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t_start).count();
    return ASETS_OK;
}
