// Device build of the particle octree: node for node the tree build_bh_tree
// makes on the host (vp_tree.h: octree::addData, octree.h:151-186, with the
// pseudo-particles of particleFMM::computeMultExp, particleFMM.cpp:29-58),
// in the same flattened pre-order layout.
//
//   1. bounding box -> root centre and half width 0.51 x the largest extent;
//   2. per particle, its octant at levels 2..22 by the host's own centre
//      recurrence (o += h * (+-0.5), h *= 0.5: the same FP64 operations, so
//      the same comparisons), packed 3 bits a level into a 63-bit key;
//   3. a stable radix sort of (key, particle index): every node is then a
//      contiguous range, leaves in pre-order;
//   4. nodes level by level: a node of more than max_members particles splits
//      into the runs of its range that share the next octant (a flag per
//      position, an exclusive scan, the children numbered in octant order);
//   5. subtree sizes bottom-up, pre-order indices and skip links top-down;
//      each leaf's members back into particle order (the host's stable
//      partition keeps insertion order inside a leaf);
//   6. pseudo-particles bottom-up (|alpha|-weighted position and summed
//      strength of the members or children, in the host's order; the root
//      has none), radius |half| sqrt(3) x 2 x 1.5.
// A node at level 22 that would still split (particles closer than 2^-21 of
// the root width) is beyond the keys: the build reports it and the caller
// builds on the host instead.
#include <hipcub/hipcub.hpp>

#include "capi_internal.h"

namespace cpgpu {

constexpr int kB = 256;

namespace capi {
namespace {

#pragma clang fp contract(off)

constexpr int kKeyLevels = 21;          // octant levels 2..22 in the key
constexpr int kDeepest = 1 + kKeyLevels;  // deepest level a key resolves

inline unsigned nblk(int64_t n) { return (unsigned)((n + kB - 1) / kB); }

__device__ inline int octant_at(uint64_t key, int level) {  // level 2..22
    return (int)((key >> (3 * (kDeepest - level))) & 7);
}

// per-block min and max of each coordinate (exact in any order)
__global__ __launch_bounds__(kB) void bbox_part_kernel(int n, const double* __restrict__ xyz,
                                                       double* __restrict__ part) {
    __shared__ double sh[6][kB];
    const int i = blockIdx.x * kB + threadIdx.x;
    double mn[3], mx[3];
    const int src = i < n ? i : 0;
    for (int d = 0; d < 3; d++) mn[d] = mx[d] = xyz[3 * src + d];
    for (int d = 0; d < 3; d++) {
        sh[d][threadIdx.x] = mn[d];
        sh[3 + d][threadIdx.x] = mx[d];
    }
    __syncthreads();
    for (int w = kB / 2; w >= 1; w >>= 1) {
        if ((int)threadIdx.x < w)
            for (int d = 0; d < 3; d++) {
                sh[d][threadIdx.x] = fmin(sh[d][threadIdx.x], sh[d][threadIdx.x + w]);
                sh[3 + d][threadIdx.x] = fmax(sh[3 + d][threadIdx.x], sh[3 + d][threadIdx.x + w]);
            }
        __syncthreads();
    }
    if (threadIdx.x < 6) part[6 * blockIdx.x + threadIdx.x] = sh[threadIdx.x][0];
}

// box[0..2] = root centre, box[3] = half width (octree::setDimensions,
// octree.h:70-80), from the per-block extremes (one workgroup)
__global__ __launch_bounds__(kB) void bbox_final_kernel(int nb, const double* __restrict__ part,
                                                        double* __restrict__ box) {
    __shared__ double sh[6][kB];
    double v[6];
    for (int d = 0; d < 6; d++) v[d] = part[d];
    for (int b = threadIdx.x; b < nb; b += kB)
        for (int d = 0; d < 3; d++) {
            v[d] = fmin(v[d], part[6 * b + d]);
            v[3 + d] = fmax(v[3 + d], part[6 * b + 3 + d]);
        }
    for (int d = 0; d < 6; d++) sh[d][threadIdx.x] = v[d];
    __syncthreads();
    for (int w = kB / 2; w >= 1; w >>= 1) {
        if ((int)threadIdx.x < w)
            for (int d = 0; d < 3; d++) {
                sh[d][threadIdx.x] = fmin(sh[d][threadIdx.x], sh[d][threadIdx.x + w]);
                sh[3 + d][threadIdx.x] = fmax(sh[3 + d][threadIdx.x], sh[3 + d][threadIdx.x + w]);
            }
        __syncthreads();
    }
    if (threadIdx.x != 0) return;
    double maxD = 0;
    for (int d = 0; d < 3; d++) {
        const double mn = sh[d][0], mx = sh[3 + d][0];
        box[d] = 0.5 * (mn + mx);
        const double e = mx - mn;
        maxD = d == 0 ? e : (e > maxD ? e : maxD);
    }
    box[3] = 0.51 * maxD;
}

__global__ __launch_bounds__(kB) void keys_kernel(int n, const double* __restrict__ xyz,
                                                  const double* __restrict__ box, uint64_t* __restrict__ keys,
                                                  int* __restrict__ idx) {
    const int i = blockIdx.x * kB + threadIdx.x;
    if (i >= n) return;
    const double x = xyz[3 * i], y = xyz[3 * i + 1], z = xyz[3 * i + 2];
    double o0 = box[0], o1 = box[1], o2 = box[2], h = box[3];
    uint64_t k = 0;
    for (int L = 0; L < kKeyLevels; L++) {
        int c = 0;
        if (x > o0) c |= 4;
        if (y > o1) c |= 2;
        if (z > o2) c |= 1;
        k = (k << 3) | (uint64_t)c;
        o0 += h * (c & 4 ? 0.5 : -0.5);
        o1 += h * (c & 2 ? 0.5 : -0.5);
        o2 += h * (c & 1 ? 0.5 : -0.5);
        h = 0.5 * h;
    }
    keys[i] = k;
    idx[i] = i;
}

struct Nodes {
    int* start;
    int* end;
    int* parent;
    int* first;   // first child (children are consecutive ids)
    int* nch;     // children
    int* size;    // subtree size
    int* pre;     // pre-order index
};

// flag[i] = 1 where a child of i's node starts (i's node at `level` splits)
__global__ __launch_bounds__(kB) void split_flags_kernel(int n, const uint64_t* __restrict__ keys,
                                                         const int* __restrict__ node_of, Nodes N, int maxm,
                                                         int level, int* __restrict__ flag, int* __restrict__ ovf) {
    const int i = blockIdx.x * kB + threadIdx.x;
    if (i >= n) return;
    const int nd = node_of[i];
    int f = 0;
    if (nd >= 0 && N.end[nd] - N.start[nd] > maxm) {
        if (level >= kDeepest) {
            *ovf = 1;
        } else {
            f = i == N.start[nd] || octant_at(keys[i], level + 1) != octant_at(keys[i - 1], level + 1);
        }
    }
    flag[i] = f;
}

// positions of splitting nodes move to their child (ids from `base` on);
// the others are done
__global__ __launch_bounds__(kB) void split_children_kernel(int n, int* __restrict__ node_of, Nodes N, int maxm,
                                                            const int* __restrict__ flag, const int* __restrict__ scan,
                                                            int base) {
    const int i = blockIdx.x * kB + threadIdx.x;
    if (i >= n) return;
    const int nd = node_of[i];
    if (nd < 0) return;
    if (N.end[nd] - N.start[nd] <= maxm) {
        node_of[i] = -1;
        return;
    }
    const int ch = base + scan[i] + flag[i] - 1;
    node_of[i] = ch;
    if (flag[i]) {
        N.start[ch] = i;
        N.parent[ch] = nd;
        if (i == N.start[nd]) N.first[nd] = ch;
        atomicAdd(&N.nch[nd], 1);
    }
}

// a child's range ends where its next sibling starts, or with its parent
__global__ __launch_bounds__(kB) void child_ends_kernel(int base, int count, Nodes N) {
    const int k = blockIdx.x * kB + threadIdx.x;
    if (k >= count) return;
    const int c = base + k;
    const int p = N.parent[c];
    N.end[c] = (k + 1 < count && N.parent[c + 1] == p) ? N.start[c + 1] : N.end[p];
    N.nch[c] = 0;
    N.first[c] = -1;
}

__global__ __launch_bounds__(kB) void subtree_size_kernel(int base, int count, Nodes N) {
    const int k = blockIdx.x * kB + threadIdx.x;
    if (k >= count) return;
    const int nd = base + k;
    int s = 1;
    for (int c = 0; c < N.nch[nd]; c++) s += N.size[N.first[nd] + c];
    N.size[nd] = s;
}

__global__ __launch_bounds__(kB) void preorder_kernel(int base, int count, Nodes N) {
    const int k = blockIdx.x * kB + threadIdx.x;
    if (k >= count) return;
    const int nd = base + k;
    int p = N.pre[nd] + 1;
    for (int c = 0; c < N.nch[nd]; c++) {
        const int ch = N.first[nd] + c;
        N.pre[ch] = p;
        p += N.size[ch];
    }
}

// a leaf's members back into particle order (insertion sort of <= a few
// dozen indices); slot of each particle
__global__ __launch_bounds__(kB) void leaf_order_kernel(int count, Nodes N, int* __restrict__ idx) {
    const int nd = blockIdx.x * kB + threadIdx.x;
    if (nd >= count || N.nch[nd] != 0) return;
    const int s0 = N.start[nd], s1 = N.end[nd];
    for (int a = s0 + 1; a < s1; a++) {
        const int v = idx[a];
        int b = a - 1;
        while (b >= s0 && idx[b] > v) {
            idx[b + 1] = idx[b];
            b--;
        }
        idx[b + 1] = v;
    }
}

__global__ __launch_bounds__(kB) void members_kernel(int n, const int* __restrict__ idx,
                                                     const double* __restrict__ xyz, const double* __restrict__ alpha,
                                                     const double* __restrict__ rad, double* __restrict__ mpos,
                                                     double* __restrict__ mstr, double* __restrict__ mrad,
                                                     int* __restrict__ slot) {
    const int s = blockIdx.x * kB + threadIdx.x;
    if (s >= n) return;
    const int i = idx[s];
    for (int d = 0; d < 3; d++) {
        mpos[3 * s + d] = xyz[3 * i + d];
        mstr[3 * s + d] = alpha[3 * i + d];
    }
    mrad[s] = rad[i];
    slot[i] = s;
}

// pseudo-particles of the nodes of one level (not the root), children done
__global__ __launch_bounds__(kB) void expansion_kernel(int base, int count, int level, Nodes N,
                                                       const int* __restrict__ idx, const double* __restrict__ xyz,
                                                       const double* __restrict__ alpha, const double* __restrict__ box,
                                                       double* __restrict__ pp) {
    const int k = blockIdx.x * kB + threadIdx.x;
    if (k >= count) return;
    const int nd = base + k;
    double wx = 0, wy = 0, wz = 0, tot = 0, sx = 0, sy = 0, sz = 0;
    auto add = [&](double px, double py, double pz, double ax, double ay, double az) {
        const double nrm = sqrt(ax * ax + ay * ay + az * az);
        wx += px * nrm;
        wy += py * nrm;
        wz += pz * nrm;
        tot += nrm;
    };
    if (N.nch[nd] == 0) {
        for (int s = N.start[nd]; s < N.end[nd]; s++) {
            const int i = idx[s];
            add(xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2], alpha[3 * i], alpha[3 * i + 1], alpha[3 * i + 2]);
        }
        for (int s = N.start[nd]; s < N.end[nd]; s++) {
            const int i = idx[s];
            sx += alpha[3 * i];
            sy += alpha[3 * i + 1];
            sz += alpha[3 * i + 2];
        }
    } else {
        for (int c = 0; c < N.nch[nd]; c++) {
            const double* q = pp + 8 * (N.first[nd] + c);
            add(q[0], q[1], q[2], q[4], q[5], q[6]);
        }
        for (int c = 0; c < N.nch[nd]; c++) {
            const double* q = pp + 8 * (N.first[nd] + c);
            sx += q[4];
            sy += q[5];
            sz += q[6];
        }
    }
    double h = box[3];
    for (int l = 1; l < level; l++) h = 0.5 * h;
    double* o = pp + 8 * nd;
    o[0] = wx / tot;
    o[1] = wy / tot;
    o[2] = wz / tot;
    o[3] = 2 * h;
    o[4] = sx;
    o[5] = sy;
    o[6] = sz;
    o[7] = sqrt(h * h + h * h + h * h) * 2 * 1.5;
}

// the flattened node of every built node at its pre-order index
__global__ __launch_bounds__(kB) void emit_kernel(int count, Nodes N, const double* __restrict__ pp,
                                                  const double* __restrict__ box, BhNode* __restrict__ out) {
    const int nd = blockIdx.x * kB + threadIdx.x;
    if (nd >= count) return;
    BhNode b{};
    if (nd == 0) {  // the root has no pseudo-particle
        b.width = 2 * box[3];
    } else {
        const double* q = pp + 8 * nd;
        b.ex = q[0];
        b.ey = q[1];
        b.ez = q[2];
        b.width = q[3];
        b.sx = q[4];
        b.sy = q[5];
        b.sz = q[6];
        b.radius = q[7];
    }
    const int pre = N.pre[nd];
    if (pre < 0 || pre >= count) return;  // (a permutation of 0..count-1)
    b.skip = pre + N.size[nd];
    if (N.nch[nd] == 0) {
        b.first = N.start[nd];
        b.count = N.end[nd] - N.start[nd];
    } else {
        b.first = 0;
        b.count = -1;
    }
    b.pad = 0;
    out[pre] = b;
}

__global__ void root_init_kernel(int n, Nodes N) {
    N.start[0] = 0;
    N.end[0] = n;
    N.parent[0] = -1;
    N.first[0] = -1;
    N.nch[0] = 0;
    N.pre[0] = 0;
}

__global__ __launch_bounds__(kB) void fill_int_kernel(int n, int v, int* __restrict__ a) {
    const int i = blockIdx.x * kB + threadIdx.x;
    if (i < n) a[i] = v;
}

}  // namespace

bool build_tree_device(cp_ctx* c, int n, const double* dxyz, const double* dalpha, const double* drad,
                       int max_members, hipStream_t s, cp_tree* t) {
    // bounding box
    const unsigned nb = nblk(n);
    DevBuf bpart, box;
    dalloc(c, bpart, (size_t)6 * nb * sizeof(double));
    dalloc(c, box, 4 * sizeof(double));
    hipLaunchKernelGGL(bbox_part_kernel, dim3(nb), dim3(kB), 0, s, n, dxyz, bpart.as<double>());
    hipLaunchKernelGGL(bbox_final_kernel, dim3(1), dim3(kB), 0, s, (int)nb, bpart.as<double>(), box.as<double>());
    // keys, stable sort
    DevBuf keys, keys2, idx, idx2;
    dalloc(c, keys, (size_t)n * sizeof(uint64_t));
    dalloc(c, keys2, (size_t)n * sizeof(uint64_t));
    dalloc(c, idx, (size_t)n * sizeof(int));
    dalloc(c, idx2, (size_t)n * sizeof(int));
    hipLaunchKernelGGL(keys_kernel, dim3(nb), dim3(kB), 0, s, n, dxyz, box.as<double>(), keys.as<uint64_t>(),
                       idx.as<int>());
    CP_LAUNCH_CHECK();
    size_t sort_bytes = 0, scan_bytes = 0;
    CP_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, sort_bytes, keys.as<uint64_t>(), keys2.as<uint64_t>(),
                                              idx.as<int>(), idx2.as<int>(), n, 0, 3 * kKeyLevels, s));
    CP_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, scan_bytes, idx.as<int>(), idx.as<int>(), n, s));
    DevBuf tmp;
    dalloc(c, tmp, std::max(sort_bytes, scan_bytes));
    CP_HIP(hipcub::DeviceRadixSort::SortPairs(tmp.p, sort_bytes, keys.as<uint64_t>(), keys2.as<uint64_t>(),
                                              idx.as<int>(), idx2.as<int>(), n, 0, 3 * kKeyLevels, s));
    const uint64_t* skeys = keys2.as<uint64_t>();
    int* sidx = idx2.as<int>();

    // nodes level by level (ids in level order; capacity 2n + 64, which a
    // tree of leaves with >= 1 particle each and at most one single-child
    // chain per particle cluster stays under in practice; beyond it the
    // caller builds on the host)
    const int cap = 2 * n + 64;
    DevBuf nb_start, nb_end, nb_parent, nb_first, nb_nch, nb_size, nb_pre, node_of, flag, scan, info;
    for (DevBuf* b : {&nb_start, &nb_end, &nb_parent, &nb_first, &nb_nch, &nb_size, &nb_pre})
        dalloc(c, *b, (size_t)cap * sizeof(int));
    dalloc(c, node_of, (size_t)n * sizeof(int));
    dalloc(c, flag, (size_t)n * sizeof(int));
    dalloc(c, scan, (size_t)n * sizeof(int));
    dalloc(c, info, 4 * sizeof(int));
    Nodes N{nb_start.as<int>(), nb_end.as<int>(), nb_parent.as<int>(), nb_first.as<int>(),
            nb_nch.as<int>(),   nb_size.as<int>(), nb_pre.as<int>()};
    hipLaunchKernelGGL(root_init_kernel, dim3(1), dim3(1), 0, s, n, N);
    hipLaunchKernelGGL(fill_int_kernel, dim3(nb), dim3(kB), 0, s, n, 0, node_of.as<int>());
    hipLaunchKernelGGL(fill_int_kernel, dim3(1), dim3(kB), 0, s, 4, 0, info.as<int>());
    std::vector<int> lvl_base{0, 1};  // level L's nodes: [lvl_base[L-1], lvl_base[L])
    int level = 1;
    for (;;) {
        const int base = lvl_base.back();
        hipLaunchKernelGGL(split_flags_kernel, dim3(nb), dim3(kB), 0, s, n, skeys, (const int*)node_of.as<int>(), N,
                           max_members, level, flag.as<int>(), info.as<int>());
        CP_HIP(hipcub::DeviceScan::ExclusiveSum(tmp.p, scan_bytes, flag.as<int>(), scan.as<int>(), n, s));
        int h[3] = {0, 0, 0};
        CP_HIP(hipMemcpyAsync(&h[0], scan.as<int>() + (n - 1), sizeof(int), hipMemcpyDeviceToHost, s));
        CP_HIP(hipMemcpyAsync(&h[1], flag.as<int>() + (n - 1), sizeof(int), hipMemcpyDeviceToHost, s));
        CP_HIP(hipMemcpyAsync(&h[2], info.as<int>(), sizeof(int), hipMemcpyDeviceToHost, s));
        CP_HIP(hipStreamSynchronize(s));
        const int total = h[0] + h[1];
        if (h[2] || base + total > cap) return false;
        if (total == 0) break;
        hipLaunchKernelGGL(split_children_kernel, dim3(nb), dim3(kB), 0, s, n, node_of.as<int>(), N, max_members,
                           (const int*)flag.as<int>(), (const int*)scan.as<int>(), base);
        hipLaunchKernelGGL(child_ends_kernel, dim3(nblk(total)), dim3(kB), 0, s, base, total, N);
        CP_LAUNCH_CHECK();
        lvl_base.push_back(base + total);
        level++;
    }
    const int nn = lvl_base.back();
    const int levels = (int)lvl_base.size() - 1;
    // sizes bottom-up, pre-order top-down
    for (int L = levels; L >= 1; L--) {
        const int b0 = lvl_base[L - 1], cnt = lvl_base[L] - b0;
        hipLaunchKernelGGL(subtree_size_kernel, dim3(nblk(cnt)), dim3(kB), 0, s, b0, cnt, N);
    }
    for (int L = 1; L < levels; L++) {
        const int b0 = lvl_base[L - 1], cnt = lvl_base[L] - b0;
        hipLaunchKernelGGL(preorder_kernel, dim3(nblk(cnt)), dim3(kB), 0, s, b0, cnt, N);
    }
    hipLaunchKernelGGL(leaf_order_kernel, dim3(nblk(nn)), dim3(kB), 0, s, nn, N, sidx);
    CP_LAUNCH_CHECK();
    // members in slot order
    dalloc(c, t->mpos, (size_t)3 * n * sizeof(double));
    dalloc(c, t->mstr, (size_t)3 * n * sizeof(double));
    dalloc(c, t->mrad, (size_t)n * sizeof(double));
    dalloc(c, t->order, (size_t)n * sizeof(int));
    dalloc(c, t->slot, (size_t)n * sizeof(int));
    hipLaunchKernelGGL(members_kernel, dim3(nb), dim3(kB), 0, s, n, (const int*)sidx, dxyz, dalpha, drad,
                       t->mpos.as<double>(), t->mstr.as<double>(), t->mrad.as<double>(), t->slot.as<int>());
    CP_HIP(hipMemcpyAsync(t->order.p, sidx, (size_t)n * sizeof(int), hipMemcpyDeviceToDevice, s));
    // pseudo-particles bottom-up (8 doubles per node), then the flat nodes
    DevBuf pp;
    dalloc(c, pp, (size_t)8 * nn * sizeof(double));
    for (int L = levels; L >= 2; L--) {
        const int b0 = lvl_base[L - 1], cnt = lvl_base[L] - b0;
        hipLaunchKernelGGL(expansion_kernel, dim3(nblk(cnt)), dim3(kB), 0, s, b0, cnt, L, N, (const int*)sidx, dxyz,
                           dalpha, (const double*)box.as<double>(), pp.as<double>());
    }
    dalloc(c, t->nodes, (size_t)nn * sizeof(BhNode));
    hipLaunchKernelGGL(emit_kernel, dim3(nblk(nn)), dim3(kB), 0, s, nn, N, (const double*)pp.as<double>(),
                       (const double*)box.as<double>(), t->nodes.as<BhNode>());
    CP_LAUNCH_CHECK();
    double hb[4];
    CP_HIP(hipMemcpyAsync(hb, box.p, 4 * sizeof(double), hipMemcpyDeviceToHost, s));
    CP_HIP(hipStreamSynchronize(s));
    t->host = BhTreeHost();
    t->host.levels = levels;
    for (int d = 0; d < 3; d++) t->host.origin[d] = hb[d];
    t->host.half = hb[3];
    t->dev.n = n;
    t->dev.n_nodes = nn;
    t->dev.nodes = t->nodes.as<BhNode>();
    t->dev.mpos = t->mpos.as<double>();
    t->dev.mstr = t->mstr.as<double>();
    t->dev.mrad = t->mrad.as<double>();
    t->dev.order = t->order.as<int>();
    return true;
}

}  // namespace capi

// ---- spatial order of a point set (the panel sums' near pass) -------------
namespace {

constexpr size_t kAlign = 256;
inline size_t align_up(size_t b) { return (b + kAlign - 1) / kAlign * kAlign; }

__device__ inline uint32_t spread10(uint32_t v) {  // 10 bits -> every third bit
    v &= 0x3ff;
    v = (v | (v << 16)) & 0x030000ffu;
    v = (v | (v << 8)) & 0x0300f00fu;
    v = (v | (v << 4)) & 0x030c30c3u;
    v = (v | (v << 2)) & 0x09249249u;
    return v;
}

// 30-bit Morton key of each point over the cube of the points' bounding box
__global__ __launch_bounds__(kB) void morton_kernel(int m, const double* __restrict__ pts,
                                                    const double* __restrict__ box, uint32_t* __restrict__ keys,
                                                    int* __restrict__ idx) {
    const int i = blockIdx.x * kB + threadIdx.x;
    if (i >= m) return;
    const double h = box[3];
    uint32_t q[3] = {0, 0, 0};
    if (h > 0)
        for (int d = 0; d < 3; d++) {
            const double u = (pts[3 * i + d] - (box[d] - h)) / (2 * h) * 1024.0;
            q[d] = (uint32_t)fmin(fmax(u, 0.0), 1023.0);
        }
    keys[i] = (spread10(q[0]) << 2) | (spread10(q[1]) << 1) | spread10(q[2]);
    idx[i] = i;
}

struct SortLayout {
    size_t parts, box, keys, keys2, idx, tmp, sort_bytes, total;
};

SortLayout sort_layout(int m) {
    SortLayout L;
    L.sort_bytes = 0;
    (void)hipcub::DeviceRadixSort::SortPairs(nullptr, L.sort_bytes, (const uint32_t*)nullptr, (uint32_t*)nullptr,
                                             (const int*)nullptr, (int*)nullptr, m, 0, 30);
    const size_t nb = capi::nblk(m);
    L.parts = 0;
    L.box = align_up(L.parts + 6 * nb * sizeof(double));
    L.keys = align_up(L.box + 4 * sizeof(double));
    L.keys2 = align_up(L.keys + (size_t)m * sizeof(uint32_t));
    L.idx = align_up(L.keys2 + (size_t)m * sizeof(uint32_t));
    L.tmp = align_up(L.idx + (size_t)m * sizeof(int));
    L.total = align_up(L.tmp + L.sort_bytes);
    return L;
}

}  // namespace

size_t morton_sort_bytes(int m) { return m > 0 ? sort_layout(m).total : 0; }

void sort_points_morton(int m, const double* pts, void* scratch, int* perm, hipStream_t s) {
    if (m <= 0) return;
    const SortLayout L = sort_layout(m);
    char* base = static_cast<char*>(scratch);
    double* parts = reinterpret_cast<double*>(base + L.parts);
    double* box = reinterpret_cast<double*>(base + L.box);
    uint32_t* keys = reinterpret_cast<uint32_t*>(base + L.keys);
    uint32_t* keys2 = reinterpret_cast<uint32_t*>(base + L.keys2);
    int* idx = reinterpret_cast<int*>(base + L.idx);
    const unsigned nb = capi::nblk(m);
    hipLaunchKernelGGL(capi::bbox_part_kernel, dim3(nb), dim3(kB), 0, s, m, pts, parts);
    hipLaunchKernelGGL(capi::bbox_final_kernel, dim3(1), dim3(kB), 0, s, (int)nb, (const double*)parts, box);
    hipLaunchKernelGGL(morton_kernel, dim3(nb), dim3(kB), 0, s, m, pts, (const double*)box, keys, idx);
    CP_LAUNCH_CHECK();
    size_t bytes = L.sort_bytes;
    CP_HIP(hipcub::DeviceRadixSort::SortPairs(base + L.tmp, bytes, keys, keys2, idx, perm, m, 0, 30, s));
}

}  // namespace cpgpu
