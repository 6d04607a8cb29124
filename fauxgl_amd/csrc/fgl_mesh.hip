// fgl_mesh.hip -- device-side mesh preparation on a resident mesh:
// BoundingBox (mesh.go:153-165), Transform (mesh.go:167-175) and the
// position-keyed normal smoothing of SmoothNormals / SmoothNormalsThreshold
// (mesh.go:80-125), with the reference's arithmetic order so that results
// are bit-identical to it.
//
// Positions and normals are the mesh's vertex arrays: vertex v at
// [3v, 3v + 3), triangles' vertices first (3 per triangle), lines' after.
//
// Smoothing groups vertices by exact position equality (Go map keys: -0 and
// +0 are one key, a NaN key matches nothing, not even itself).  Vertices are
// stably radix-sorted by a 32-bit hash of the position, so equal positions
// are contiguous and in vertex order (the order the reference appends them
// to the key's list); each vertex then walks its hash run, keeps the members
// whose position equals its own, and sums their normals in that order.
#include "fgl_internal.h"

namespace fgl {

// ---------------------------------------------------------------------------
// BoundingBox: Box.Extend (box.go:48-53) replaces the running box while it
// equals EmptyBox (all six fields == 0), so primitives before the first one
// with a coordinate != 0 do not count; after it, min / max of every
// coordinate (math.Min / Max, order-independent).

// first vertex index (in primitive order) of a primitive with a nonzero
// coordinate: per-block minimum, written without atomics
__global__ __launch_bounds__(256) void k_mesh_first_nonzero(const double *pos, int64_t nprim, int32_t nv,
                                                            int64_t *block_first) {
    __shared__ int64_t s[256];
    int64_t first = INT64_MAX;
    for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < nprim; p += (int64_t)gridDim.x * blockDim.x) {
        const double *q = pos + 3 * nv * p;
        bool nz = false;
        for (int k = 0; k < 3 * nv; k++) nz |= !(q[k] == 0.0);
        if (nz && p < first) first = p;
    }
    s[threadIdx.x] = first;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if (threadIdx.x < off && s[threadIdx.x + off] < s[threadIdx.x]) s[threadIdx.x] = s[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) block_first[blockIdx.x] = s[0];
}

// per-block min / max over the vertices of primitives p >= from
__global__ __launch_bounds__(256) void k_mesh_box(const double *pos, int64_t nprim, int32_t nv, int64_t from,
                                                  double *block_box) {
    __shared__ double s[6][256];
    double lo[3] = {__builtin_inf(), __builtin_inf(), __builtin_inf()};
    double hi[3] = {-__builtin_inf(), -__builtin_inf(), -__builtin_inf()};
    bool any = false;
    const int64_t nvert = nprim * nv, v0 = from * nv;
    for (int64_t v = v0 + (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < nvert;
         v += (int64_t)gridDim.x * blockDim.x) {
        for (int k = 0; k < 3; k++) {
            const double x = pos[3 * v + k];
            lo[k] = any ? go_min(lo[k], x) : x;
            hi[k] = any ? go_max(hi[k], x) : x;
        }
        any = true;
    }
    // threads without a vertex contribute +Inf / -Inf, the identities of
    // math.Min / Max except against NaN (which they leave to the others)
    for (int k = 0; k < 3; k++) {
        s[k][threadIdx.x] = lo[k];
        s[3 + k][threadIdx.x] = hi[k];
    }
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if (threadIdx.x < off)
            for (int k = 0; k < 3; k++) {
                s[k][threadIdx.x] = go_min(s[k][threadIdx.x], s[k][threadIdx.x + off]);
                s[3 + k][threadIdx.x] = go_max(s[3 + k][threadIdx.x], s[3 + k][threadIdx.x + off]);
            }
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int k = 0; k < 6; k++) block_box[6 * blockIdx.x + k] = s[k][0];
}

// ---------------------------------------------------------------------------
// Transform (triangle.go:66-73, line.go:23-28): MulPosition and MulDirection
// (matrix.go:209-229), the latter normalized (vector.go:83-86)

struct Mat4 {
    double m[16];
};

__global__ __launch_bounds__(256) void k_mesh_transform(double *pos, double *nrm, int64_t nvert, Mat4 M) {
    const double *a = M.m;
    for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < nvert; v += (int64_t)gridDim.x * blockDim.x) {
        const double x = pos[3 * v], y = pos[3 * v + 1], z = pos[3 * v + 2];
        pos[3 * v] = a[0] * x + a[1] * y + a[2] * z + a[3];
        pos[3 * v + 1] = a[4] * x + a[5] * y + a[6] * z + a[7];
        pos[3 * v + 2] = a[8] * x + a[9] * y + a[10] * z + a[11];
        const double nx = nrm[3 * v], ny = nrm[3 * v + 1], nz = nrm[3 * v + 2];
        const double dx = a[0] * nx + a[1] * ny + a[2] * nz;
        const double dy = a[4] * nx + a[5] * ny + a[6] * nz;
        const double dz = a[8] * nx + a[9] * ny + a[10] * nz;
        const double r = 1.0 / sqrt(dx * dx + dy * dy + dz * dz);
        nrm[3 * v] = dx * r;
        nrm[3 * v + 1] = dy * r;
        nrm[3 * v + 2] = dz * r;
    }
}

// ---------------------------------------------------------------------------
// grouping by position

FGL_D uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

// 32-bit hash of the position with -0 folded onto +0; vals = vertex index
__global__ __launch_bounds__(256) void k_mesh_hash(const double *pos, int64_t nvert, uint32_t *keys, uint32_t *vals) {
    for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < nvert; v += (int64_t)gridDim.x * blockDim.x) {
        uint64_t h = 0x9e3779b97f4a7c15ull;
        for (int k = 0; k < 3; k++) {
            const double x = pos[3 * v + k] + 0.0;  // -0 + 0 = +0
            h = mix64(h ^ (uint64_t)__double_as_longlong(x));
        }
        keys[v] = (uint32_t)(h >> 32);
        vals[v] = (uint32_t)v;
    }
}

// stable LSD radix sort of (key, val), 8 bits per pass (as in fgl_bin.hip,
// with a host-known count)
constexpr int MS_THREADS = 256, MS_ITEMS = 16, MS_BLOCK = MS_THREADS * MS_ITEMS;

__global__ __launch_bounds__(MS_THREADS) void k_ms_hist(const uint32_t *keys, uint32_t n, int shift, uint32_t *hist,
                                                        int32_t nblocks) {
    __shared__ uint32_t h[256];
    h[threadIdx.x] = 0;
    __syncthreads();
    const uint64_t base = (uint64_t)blockIdx.x * MS_BLOCK;
    uint32_t k[MS_ITEMS];
#pragma unroll
    for (int it = 0; it < MS_ITEMS; it++) {
        const uint64_t idx = base + (uint64_t)it * MS_THREADS + threadIdx.x;
        k[it] = keys[idx < n ? idx : n - 1];
    }
#pragma unroll
    for (int it = 0; it < MS_ITEMS; it++) {
        const uint64_t idx = base + (uint64_t)it * MS_THREADS + threadIdx.x;
        if (idx < n) atomicAdd(&h[(k[it] >> shift) & 255u], 1u);
    }
    __syncthreads();
    hist[(uint64_t)threadIdx.x * nblocks + blockIdx.x] = h[threadIdx.x];
}

FGL_D uint32_t ms_block_scan(uint32_t v, uint32_t *wsum, uint32_t *total) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint32_t x = v;
    for (int off = 1; off < 64; off <<= 1) {
        const uint32_t y = __shfl_up(x, off);
        if (lane >= off) x += y;
    }
    if (lane == 63) wsum[wave] = x;
    __syncthreads();
    uint32_t base = 0, tot = 0;
    for (int w = 0; w < MS_THREADS / 64; w++) {
        if (w < wave) base += wsum[w];
        tot += wsum[w];
    }
    __syncthreads();
    *total = tot;
    return base + x - v;
}

// one workgroup per digit: exclusive scan of its per-block counts
__global__ __launch_bounds__(MS_THREADS) void k_ms_scan(uint32_t *hist, int32_t nblocks, uint32_t *dtot) {
    __shared__ uint32_t wsum[MS_THREADS / 64];
    uint32_t *row = hist + (uint64_t)blockIdx.x * nblocks;
    uint32_t carry = 0;
    for (int base = 0; base < nblocks; base += MS_THREADS) {
        const int k = base + threadIdx.x;
        const uint32_t v = k < nblocks ? row[k] : 0u;
        uint32_t tot;
        const uint32_t ex = ms_block_scan(v, wsum, &tot);
        if (k < nblocks) row[k] = carry + ex;
        carry += tot;
    }
    if (threadIdx.x == 0) dtot[blockIdx.x] = carry;
}

__global__ __launch_bounds__(MS_THREADS) void k_ms_scatter(const uint32_t *kin, const uint32_t *vin, uint32_t n,
                                                           int shift, const uint32_t *hist, const uint32_t *dtot,
                                                           int32_t nblocks, uint32_t *kout, uint32_t *vout) {
    __shared__ uint32_t run[256];
    __shared__ uint32_t wcnt[MS_THREADS / 64][256];
    __shared__ uint32_t wsum[MS_THREADS / 64];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const uint64_t ltmask = (1ull << lane) - 1ull;
    {
        uint32_t tot;
        const uint32_t dbase = ms_block_scan(dtot[tid], wsum, &tot);
        run[tid] = dbase + hist[(uint64_t)tid * nblocks + blockIdx.x];
    }
    const uint64_t base = (uint64_t)blockIdx.x * MS_BLOCK;
    uint32_t kk[MS_ITEMS], vv[MS_ITEMS];
#pragma unroll
    for (int it = 0; it < MS_ITEMS; it++) {
        const uint64_t idx = base + (uint64_t)it * MS_THREADS + tid;
        const uint64_t ci = idx < n ? idx : n - 1;
        kk[it] = kin[ci];
        vv[it] = vin[ci];
    }
#pragma unroll
    for (int it = 0; it < MS_ITEMS; it++) {
        for (int w = 0; w < MS_THREADS / 64; w++) wcnt[w][tid] = 0;
        __syncthreads();
        const uint64_t idx = base + (uint64_t)it * MS_THREADS + tid;
        const bool valid = idx < n;
        const uint32_t key = valid ? kk[it] : 0u;
        const uint32_t digit = (key >> shift) & 255u;
        uint64_t peers = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; b++) {
            const bool bit = (digit >> b) & 1u;
            const uint64_t m = __ballot(bit);
            peers &= bit ? m : ~m;
        }
        const uint32_t rank = (uint32_t)__popcll(peers & ltmask);
        if (valid && rank == 0) wcnt[wave][digit] = (uint32_t)__popcll(peers);
        __syncthreads();
        {
            uint32_t r = run[tid];
            for (int w = 0; w < MS_THREADS / 64; w++) {
                const uint32_t c = wcnt[w][tid];
                wcnt[w][tid] = r;
                r += c;
            }
            run[tid] = r;
        }
        __syncthreads();
        if (valid) {
            const uint32_t p = wcnt[wave][digit] + rank;
            kout[p] = key;
            vout[p] = vv[it];
        }
        __syncthreads();
    }
}

// where each vertex landed in the sorted order
__global__ __launch_bounds__(256) void k_mesh_rank(const uint32_t *svals, int64_t n, uint32_t *where) {
    for (int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; s < n; s += (int64_t)gridDim.x * blockDim.x)
        where[svals[s]] = (uint32_t)s;
}

FGL_D bool same_key(const double *pos, uint32_t a, uint32_t b) {
    return pos[3 * a] == pos[3 * b] && pos[3 * a + 1] == pos[3 * b + 1] && pos[3 * a + 2] == pos[3 * b + 2];
}

// THRESHOLD: smoothNormalsThreshold (mesh.go:80-89) of each vertex's own
// normal against the key's list; otherwise SmoothNormals (mesh.go:105-125):
// the normalized sum of the whole list.  The sum starts at the zero vector
// and adds the members in vertex order.
template <bool THRESHOLD>
__global__ __launch_bounds__(256) void k_mesh_smooth(const double *pos, const double *nrm, int64_t nvert,
                                                     const uint32_t *skeys, const uint32_t *svals,
                                                     const uint32_t *where, double threshold, double *out) {
    for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < nvert; v += (int64_t)gridDim.x * blockDim.x) {
        const uint32_t s = where[v], h = skeys[s];
        int64_t a = s;
        while (a > 0 && skeys[a - 1] == h) a--;
        const double ox = nrm[3 * v], oy = nrm[3 * v + 1], oz = nrm[3 * v + 2];
        double rx = 0.0, ry = 0.0, rz = 0.0;
        bool listed = false;  // the key has a list (a NaN position has none)
        for (int64_t j = a; j < nvert && skeys[j] == h; j++) {
            const uint32_t u = svals[j];
            if (!same_key(pos, u, (uint32_t)v)) continue;
            listed = true;
            const double x = nrm[3 * u], y = nrm[3 * u + 1], z = nrm[3 * u + 2];
            if (THRESHOLD && !(x * ox + y * oy + z * oz >= threshold)) continue;
            rx = rx + x;
            ry = ry + y;
            rz = rz + z;
        }
        if (!THRESHOLD && !listed) {
            // SmoothNormals reads lookup[position] for a key it never stored:
            // the zero Vector, not normalized
            out[3 * v] = 0.0;
            out[3 * v + 1] = 0.0;
            out[3 * v + 2] = 0.0;
            continue;
        }
        const double r = 1.0 / sqrt(rx * rx + ry * ry + rz * rz);
        out[3 * v] = rx * r;
        out[3 * v + 1] = ry * r;
        out[3 * v + 2] = rz * r;
    }
}

// ---------------------------------------------------------------------------

static unsigned mesh_grid(int64_t n) {
    const int64_t b = (n + 255) / 256;
    return (unsigned)(b < 1 ? 1 : (b > 8192 ? 8192 : b));
}

hipError_t launch_mesh_first_nonzero(const double *pos, int64_t nprim, int32_t nv, int64_t *block_first,
                                     int32_t nblocks, hipStream_t s) {
    hipLaunchKernelGGL(k_mesh_first_nonzero, dim3(nblocks), dim3(256), 0, s, pos, nprim, nv, block_first);
    return hipGetLastError();
}
hipError_t launch_mesh_box(const double *pos, int64_t nprim, int32_t nv, int64_t from, double *block_box,
                           int32_t nblocks, hipStream_t s) {
    hipLaunchKernelGGL(k_mesh_box, dim3(nblocks), dim3(256), 0, s, pos, nprim, nv, from, block_box);
    return hipGetLastError();
}
hipError_t launch_mesh_transform(double *pos, double *nrm, int64_t nvert, const double m[16], hipStream_t s) {
    Mat4 M;
    for (int k = 0; k < 16; k++) M.m[k] = m[k];
    hipLaunchKernelGGL(k_mesh_transform, dim3(mesh_grid(nvert)), dim3(256), 0, s, pos, nrm, nvert, M);
    return hipGetLastError();
}
int32_t mesh_sort_blocks(int64_t n) { return (int32_t)((n + MS_BLOCK - 1) / MS_BLOCK); }
hipError_t launch_mesh_group(const double *pos, int64_t nvert, uint32_t *keys[2], uint32_t *vals[2], uint32_t *hist,
                             uint32_t *dtot, uint32_t *where, hipStream_t s) {
    const int32_t nb = mesh_sort_blocks(nvert);
    hipLaunchKernelGGL(k_mesh_hash, dim3(mesh_grid(nvert)), dim3(256), 0, s, pos, nvert, keys[0], vals[0]);
    for (int pass = 0; pass < 4; pass++) {
        const int a = pass & 1, b = a ^ 1;
        hipLaunchKernelGGL(k_ms_hist, dim3(nb), dim3(MS_THREADS), 0, s, keys[a], (uint32_t)nvert, 8 * pass, hist, nb);
        hipLaunchKernelGGL(k_ms_scan, dim3(256), dim3(MS_THREADS), 0, s, hist, nb, dtot);
        hipLaunchKernelGGL(k_ms_scatter, dim3(nb), dim3(MS_THREADS), 0, s, keys[a], vals[a], (uint32_t)nvert, 8 * pass,
                           hist, dtot, nb, keys[b], vals[b]);
    }
    // four passes: the sorted pairs are back in keys[0] / vals[0]
    hipLaunchKernelGGL(k_mesh_rank, dim3(mesh_grid(nvert)), dim3(256), 0, s, vals[0], nvert, where);
    return hipGetLastError();
}
hipError_t launch_mesh_smooth(bool threshold_mode, const double *pos, const double *nrm, int64_t nvert,
                              const uint32_t *skeys, const uint32_t *svals, const uint32_t *where, double threshold,
                              double *out, hipStream_t s) {
    if (threshold_mode)
        hipLaunchKernelGGL(k_mesh_smooth<true>, dim3(mesh_grid(nvert)), dim3(256), 0, s, pos, nrm, nvert, skeys, svals,
                           where, threshold, out);
    else
        hipLaunchKernelGGL(k_mesh_smooth<false>, dim3(mesh_grid(nvert)), dim3(256), 0, s, pos, nrm, nvert, skeys,
                           svals, where, threshold, out);
    return hipGetLastError();
}

}  // namespace fgl
