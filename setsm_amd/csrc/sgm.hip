// sgm.hip — SGM 8-path cost aggregation (setsm_code.cpp:7017-7509 with
// SGM_start_pos :6755-6767 and SGM_con_pos :6769-6999) over the CSR voxel
// cube, for AWNCC at levels <= SGM_py.
//
// Built with -ffp-contract=off: the path costs are float, rounded from the
// reference's double expressions in its operand order, and SumCost receives
// one float addition per (direction pass, cell, height) in the reference's
// direction order, so the result is bit-identical to the CPU path.
//
// Decomposition: one launch per loop of the reference (12 loops: four
// straight directions with one path per row or column, and for each diagonal
// a loop of paths from a side wall and a loop from the top/bottom wall; the
// corner path is in both, as in the reference). The paths of one launch touch
// disjoint cells, so SumCost needs no atomics. One wavefront walks one path;
// its lanes cover the heights of the current cell. The previous cell's path
// costs and voxel flags sit in a ping-pong pair of buffers, in LDS when the
// longest height list fits, else in a per-path global scratch.

#include "smx_internal.hpp"

namespace smx {

struct SgmArgs {
    int W, H;
    double step;
    const smx_nccresult *ncc;
    const smx_ugrid *gridpt3;
    const int64_t *vox_off;
    const uint32_t *vox;
    float *sum;          // SumCost, CSR like the cube
    int d;               // reference direction index 0..7
    int kind;            // 0: one path per row (side wall), 1: one path per column (top / bottom wall)
    int dr, dc;          // step from one cell of the path to the next
    int cap;             // buffer capacity in heights (>= max NumOfHeight, multiple of 4)
    unsigned char *gscratch;   // per-path scratch (10 * cap bytes) when not in LDS
};

extern __shared__ unsigned char smx_sgm_lds[];

// The reference's `a > b ? a : b`.
__device__ __forceinline__ double max2(double a, double b) { return a > b ? a : b; }

template <bool kLds>
__global__ void __launch_bounds__(64) k_sgm_path(SgmArgs A)
{
    const int p = blockIdx.x;
    const int lane = threadIdx.x;
    unsigned char *base = kLds ? smx_sgm_lds : A.gscratch + (size_t)p * 10 * (size_t)A.cap;
    float *pre = (float *)base;
    float *cur = pre + A.cap;
    unsigned char *fpre = (unsigned char *)(cur + A.cap);
    unsigned char *fcur = fpre + A.cap;

    int row, col;
    if (A.kind == 0) {
        row = p;
        col = A.dc > 0 ? 0 : A.W - 1;
    } else {
        col = p;
        row = A.dr > 0 ? 0 : A.H - 1;
    }
    const double step = A.step;
    const double P1 = 0.3, P2 = 0.6;

    // SGM_start_pos
    long c = (long)row * A.W + col;
    // heights of a cell: NumOfHeight, never past the cell's row of the cube
    auto row_len = [&](long cc) {
        return (int)min((long)A.ncc[cc].NumOfHeight, (long)(A.vox_off[cc + 1] - A.vox_off[cc]));
    };
    int nh = row_len(c);
    int minH = A.ncc[c].minHeight;
    short gmin = A.gridpt3[c].minHeight, gmax = A.gridpt3[c].maxHeight;
    int64_t off = A.vox_off[c];
    for (int h = lane; h < nh; h += 64) {
        const uint32_t w = A.vox[off + h];
        const float ih = (float)(minH + h * step);
        float v = 0.0f;
        if (ih >= gmin && ih <= gmax) {
            v = (float)dq(SMX_VOXEL_INCC(w));
            A.sum[off + h] += v;
        }
        pre[h] = v;
        fpre[h] = (unsigned char)SMX_VOXEL_FLAG(w);
    }
    if (!kLds) __threadfence_block();
    __syncthreads();

    int t_nh = nh, t_minH = minH;
    float t_maxW = (float)dq(A.ncc[c].max_WNCC);
    // The next cell's record, search band and row offset are loaded one step
    // ahead: a step then waits for its cube words and sums only.
    struct Meta {
        smx_nccresult nr;
        short gmin, gmax;
        int64_t off, end;
    };
    auto load_meta = [&](long cc) {
        Meta m;
        m.nr = A.ncc[cc];
        m.gmin = A.gridpt3[cc].minHeight;
        m.gmax = A.gridpt3[cc].maxHeight;
        m.off = A.vox_off[cc];
        m.end = A.vox_off[cc + 1];
        return m;
    };
    auto inside = [&](int r, int q) { return r >= 0 && r < A.H && q >= 0 && q < A.W; };
    // The first kPre * 64 heights of the next cell's cube words and sums are
    // loaded one step ahead into registers (the cell's record two steps
    // ahead), so a step waits for LDS and its arithmetic only; a cell's sums
    // are touched by no other path of the launch in between.
    constexpr int kPre = 2;
    uint32_t wn[kPre];
    float sn[kPre];
    auto prefetch = [&](const Meta &mm) {
        const int n2 = (int)min((long)mm.nr.NumOfHeight, (long)(mm.end - mm.off));
#pragma unroll
        for (int j = 0; j < kPre; j++) {
            const int h = lane + 64 * j;
            wn[j] = h < n2 ? A.vox[mm.off + h] : 0u;
            sn[j] = h < n2 ? A.sum[mm.off + h] : 0.0f;
        }
    };
    row += A.dr;
    col += A.dc;
    Meta next{}, next2{};
    if (inside(row, col)) {
        next = load_meta((long)row * A.W + col);
        prefetch(next);
    }
    if (inside(row + A.dr, col + A.dc)) next2 = load_meta((long)(row + A.dr) * A.W + col + A.dc);
    for (; inside(row, col); row += A.dr, col += A.dc) {
        // SGM_con_pos for cell (row, col); its predecessor on the path,
        // (row + v_row[d], col + u_col[d]) in the reference's tables
        // (setsm_code.cpp:7036-7037), is the previous cell.
        c = (long)row * A.W + col;
        const Meta m = next;
        uint32_t wc[kPre];
        float sc[kPre];
#pragma unroll
        for (int j = 0; j < kPre; j++) {
            wc[j] = wn[j];
            sc[j] = sn[j];
        }
        next = next2;
        if (inside(row + A.dr, col + A.dc)) prefetch(next);
        if (inside(row + 2 * A.dr, col + 2 * A.dc))
            next2 = load_meta((long)(row + 2 * A.dr) * A.W + col + 2 * A.dc);
        nh = (int)min((long)m.nr.NumOfHeight, (long)(m.end - m.off));
        minH = m.nr.minHeight;
        gmin = m.gmin;
        gmax = m.gmax;
        off = m.off;
        const double V4 = (double)t_maxW - P2;
        // one height of the cell: its cube word w and its sum before this path
        auto visit = [&](int h, uint32_t w, float sum_before) {
            const float ih = (float)(minH + h * step);
            float out = 0.0f;
            if (ih >= gmin && ih <= gmax) {
                const double WNCC_sum = dq(SMX_VOXEL_INCC(w));
                const int i0 = (int)((double)(ih - (float)t_minH) / step);
                const int i1 = i0 - 1, i2 = i0 + 1;
                bool use1 = false, use2 = false, use3 = false, zero3 = false;
                if (i0 == 0 && i2 >= 0 && i2 < t_nh) {
                    use3 = fpre[i0] != 0;
                    use2 = fpre[i2] != 0;
                } else if (i0 == t_nh - 1 && i1 >= 0 && i1 < t_nh) {
                    use3 = fpre[i0] != 0;
                    use1 = fpre[i1] != 0;
                } else if (i0 > 0 && i0 < t_nh - 1 && i1 >= 0 && i1 < t_nh && i2 >= 0 && i2 < t_nh) {
                    use3 = fpre[i0] != 0;
                    use1 = fpre[i1] != 0;
                    use2 = fpre[i2] != 0;
                    // setsm_code.cpp:6917-6929 compares V3 = 0 in this case
                    zero3 = !use3 && use1 && use2;
                }
                double m = V4;
                if (use1) m = max2((double)pre[i1] - P1, m);
                if (use2) m = max2((double)pre[i2] - P1, m);
                if (use3) m = max2((double)pre[i0], m);
                if (zero3) m = max2(0.0, m);
                const double t_WNCC_sum = m - (double)t_maxW;
                out = (float)(WNCC_sum + t_WNCC_sum);
                A.sum[off + h] = sum_before + out;
            }
            cur[h] = out;
            fcur[h] = (unsigned char)SMX_VOXEL_FLAG(w);
        };
#pragma unroll
        for (int j = 0; j < kPre; j++)
            if (lane + 64 * j < nh) visit(lane + 64 * j, wc[j], sc[j]);
        for (int h = lane + 64 * kPre; h < nh; h += 64) visit(h, A.vox[off + h], A.sum[off + h]);
        if (!kLds) __threadfence_block();
        __syncthreads();
        float *tf = pre;
        pre = cur;
        cur = tf;
        unsigned char *tb = fpre;
        fpre = fcur;
        fcur = tb;
        t_nh = nh;
        t_minH = minH;
        t_maxW = (float)dq(m.nr.max_WNCC);
    }
}

__global__ void __launch_bounds__(256) k_max_nh(const smx_nccresult *__restrict__ ncc, long n,
                                                unsigned long long *out)
{
    const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
    unsigned v = c < n ? ncc[c].NumOfHeight : 0u;
    for (int o = 32; o > 0; o >>= 1) v = max(v, (unsigned)__shfl_xor((int)v, o));
    if ((threadIdx.x & 63) == 0 && v) atomicMax(out, (unsigned long long)v);
}

hipError_t launch_max_nh(const smx_nccresult *ncc, long n, unsigned long long *out, hipStream_t s)
{
    hipLaunchKernelGGL(k_max_nh, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, ncc, n, out);
    return hipGetLastError();
}

size_t sgm_lds_limit() { return 64 * 1024; }

// The reference's 12 loops in order (setsm_code.cpp:7056-7505).
hipError_t launch_sgm(const SgmLaunch &L, hipStream_t s)
{
    static const int loops[12][4] = {
        {0, 0, 0, 1}, {1, 0, 0, -1}, {2, 1, 1, 0}, {3, 1, -1, 0},
        {4, 0, 1, 1}, {4, 1, 1, 1}, {5, 0, 1, -1}, {5, 1, 1, -1},
        {6, 0, -1, 1}, {6, 1, -1, 1}, {7, 0, -1, -1}, {7, 1, -1, -1},
    };
    SgmArgs A{};
    A.W = L.W;
    A.H = L.H;
    A.step = L.step;
    A.ncc = L.ncc;
    A.gridpt3 = L.gridpt3;
    A.vox_off = L.vox_off;
    A.vox = L.vox;
    A.sum = L.sum;
    A.cap = L.cap;
    A.gscratch = L.gscratch;
    const size_t lds = (size_t)10 * (size_t)L.cap;
    for (int k = 0; k < 12; k++) {
        A.d = loops[k][0];
        A.kind = loops[k][1];
        A.dr = loops[k][2];
        A.dc = loops[k][3];
        const unsigned npaths = (unsigned)(A.kind == 0 ? A.H : A.W);
        if (L.gscratch)
            hipLaunchKernelGGL(k_sgm_path<false>, dim3(npaths), dim3(64), 0, s, A);
        else
            hipLaunchKernelGGL(k_sgm_path<true>, dim3(npaths), dim3(64), lds, s, A);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

}  // namespace smx
