// Interface updates of the partitioned (Dirichlet-Neumann) coupling.  Each
// expression follows the grouping of the patch-field code it restates, and
// the sums are double-double, so the results do not depend on the block
// size or the order in which faces are reduced.
#include "dd.hpp"
#include "dna.hpp"

namespace mrfg {

namespace {

constexpr int kDnaBlock = 256;
constexpr double kSmallFoam = 1e-15;   // doubleScalarSMALL

// Block sum of NV double-double accumulators; every thread gets the result.
template <int NV>
__device__ void blockSum(DD (&acc)[NV], double (&out)[NV])
{
    __shared__ DD sh[kDnaBlock / 64][NV];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
    for (int j = 0; j < NV; ++j) acc[j] = waveSum(acc[j]);
    if (lane == 0)
        for (int j = 0; j < NV; ++j) sh[wid][j] = acc[j];
    __syncthreads();
#pragma unroll
    for (int j = 0; j < NV; ++j) {
        DD t = sh[0][j];
        for (int w = 1; w < kDnaBlock / 64; ++w) t = ddAdd(t, sh[w][j]);
        out[j] = t.hi + t.lo;
    }
    __syncthreads();
}

// The relaxation factor of one call of an acceleration model
// (fixedRelaxation::relax, aitkenRelaxation::updateAitkenFactor): res / prevRes
// already hold this call's and the previous call's residual fields.  Every
// path passes a barrier after the side's state is read, so thread 0 may
// update it afterwards.
__device__ double relaxFactor(const DnaSide &S, const double *res, const double *prevRes, double *omega, int *corr,
                              int sideId)
{
    const int c = corr[sideId] == 0 ? 1 : corr[sideId];
    double w = omega[sideId];
    if (S.accType == 1) {
        if (c < 3) {
            w = w < S.relax ? S.relax : w;
            __syncthreads();
        } else {
            DD acc[2] = {DD{0.0, 0.0}, DD{0.0, 0.0}};
            for (int i = S.begin + threadIdx.x; i < S.end; i += kDnaBlock) {
                const double d = res[i] - prevRes[i];
                acc[0] = ddAdd(acc[0], DD{prevRes[i] * d, 0.0});
                acc[1] = ddAdd(acc[1], DD{d * d, 0.0});
            }
            double s[2];
            blockSum<2>(acc, s);
            w = -w * (s[0] / (s[1] + kSmallFoam));
            w = fabs(w);
            if (w > 1.0) w = 1.0;
            else if (w < kSmallFoam) w = S.relax;
        }
    } else {
        w = S.relax;
        __syncthreads();
    }
    return w;
}

__device__ void finishRelax(const DnaSide &S, double w, double *val, double *prev, const double *res,
                            double *omega, int *corr, int sideId)
{
    for (int i = S.begin + threadIdx.x; i < S.end; i += kDnaBlock) {
        const double out = prev[i] + w * res[i];
        val[i] = out;
        prev[i] = out;
    }
    if (threadIdx.x == 0) {
        if (S.accType == 1) omega[sideId] = w;
        corr[sideId] = (corr[sideId] == 0 ? 1 : corr[sideId]) + 1;
    }
}

// IQN-ILS (IQNILS.C:70-170, IQNILSTemplates.C:37-125) on one side: val
// holds this call's field before the update, res its residual.  The second
// correction stores the references; from the third on the residual and
// field differences are appended; with two columns or more the update is
// val + sum_t W_t c_(cols-1-t), c from a modified Gram-Schmidt QR of the
// columns (newest first) and a triangular solve that drops columns whose
// pivot is below 1e-10 times the largest column sum; else fixed relaxation.
__device__ void iqnils(const DnaSide &S, double *val, double *prev, const double *res, double *hist, int *cols,
                       int *corr, int sideId)
{
    const int c = corr[sideId] == 0 ? 1 : corr[sideId];
    int ncol = cols[sideId];
    __syncthreads();   // the side's state is read before thread 0 updates it
    const int n = S.end - S.begin, mc = S.maxCols;
    double *V = hist + S.histBase, *W = V + (size_t)mc * n, *Q = W + (size_t)mc * n;
    double *fRef = Q + (size_t)mc * n, *rRef = fRef + n, *R = rRef + n, *Cv = R + (size_t)mc * mc;
    if (c == 2) {
        for (int i = S.begin + threadIdx.x; i < S.end; i += kDnaBlock) {
            fRef[i - S.begin] = val[i];
            rRef[i - S.begin] = res[i];
        }
    } else if (c > 2 && ncol < mc) {
        for (int i = S.begin + threadIdx.x; i < S.end; i += kDnaBlock) {
            const int k = i - S.begin;
            V[(size_t)ncol * n + k] = res[i] - rRef[k];
            W[(size_t)ncol * n + k] = val[i] - fRef[k];
        }
        ++ncol;
    }
    if (ncol > 1) {
        for (int i = S.begin + threadIdx.x; i < S.end; i += kDnaBlock) {
            const int k = i - S.begin;
            for (int t = 0; t < ncol; ++t) Q[(size_t)t * n + k] = V[(size_t)(ncol - 1 - t) * n + k];
        }
        for (int a = 0; a < ncol; ++a) {
            double *Qa = Q + (size_t)a * n;
            DD acc[1] = {DD{0.0, 0.0}};
            for (int k = threadIdx.x; k < n; k += kDnaBlock) acc[0] = ddAdd(acc[0], DD{Qa[k] * Qa[k], 0.0});
            double s[1];
            blockSum<1>(acc, s);
            const double raa = sqrt(s[0]);
            if (threadIdx.x == 0) R[(size_t)a * mc + a] = raa;
            for (int k = threadIdx.x; k < n; k += kDnaBlock) Qa[k] = Qa[k] / raa;
            for (int b = a + 1; b < ncol; ++b) {
                double *Qb = Q + (size_t)b * n;
                DD d[1] = {DD{0.0, 0.0}};
                for (int k = threadIdx.x; k < n; k += kDnaBlock) d[0] = ddAdd(d[0], DD{Qa[k] * Qb[k], 0.0});
                double r[1];
                blockSum<1>(d, r);
                if (threadIdx.x == 0) R[(size_t)a * mc + b] = r[0];
                for (int k = threadIdx.x; k < n; k += kDnaBlock) Qb[k] = Qb[k] - r[0] * Qa[k];
            }
            DD e[1] = {DD{0.0, 0.0}};
            for (int k = threadIdx.x; k < n; k += kDnaBlock) e[0] = ddAdd(e[0], DD{Qa[k] * (-res[S.begin + k]), 0.0});
            double cc[1];
            blockSum<1>(e, cc);
            if (threadIdx.x == 0) Cv[a] = cc[0];
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            double eps = 0.0;
            for (int j = 0; j < ncol; ++j) {
                double t = 0.0;
                for (int i = 0; i < j + 1; ++i) t += fabs(R[(size_t)i * mc + j]);
                if (j == 0 || t > eps) eps = t;
            }
            eps = 1.0e-10 * eps;
            for (int i = 0; i < ncol; ++i)
                if (fabs(R[(size_t)i * mc + i]) > eps) {
                    for (int j = i + 1; j < ncol; ++j) R[(size_t)i * mc + j] = R[(size_t)i * mc + j] / R[(size_t)i * mc + i];
                    Cv[i] = Cv[i] / R[(size_t)i * mc + i];
                    R[(size_t)i * mc + i] = 1.0;
                }
            for (int j = ncol - 1; j >= 0; --j) {
                if (fabs(R[(size_t)j * mc + j]) > eps) {
                    for (int i = 0; i < j; ++i) Cv[i] = Cv[i] - Cv[j] * R[(size_t)i * mc + j];
                } else {
                    Cv[j] = 0.0;
                }
            }
        }
        __syncthreads();
        for (int i = S.begin + threadIdx.x; i < S.end; i += kDnaBlock) {
            const int k = i - S.begin;
            double out = val[i];
            for (int t = 0; t < ncol; ++t) out = out + W[(size_t)t * n + k] * Cv[ncol - 1 - t];
            val[i] = out;
            prev[i] = out;
        }
    } else {
        for (int i = S.begin + threadIdx.x; i < S.end; i += kDnaBlock) {
            const double out = prev[i] + S.relax * res[i];
            val[i] = out;
            prev[i] = out;
        }
    }
    if (threadIdx.x == 0) {
        cols[sideId] = ncol;
        corr[sideId] = c + 1;
    }
}

// k snGrad T of Dirichlet face i: ((T_b - T_P) deltaCoeffs) k
__device__ inline double dirFlux(const DnaDev &m, int i)
{
    return ((m.dT[i] - m.x[m.dPos[i]]) * m.dDc[i]) * m.dKb[i];
}

// Dirichlet side: T* = nbr + nbr (K - 1) from the Neumann patch values, relaxed
__global__ __launch_bounds__(kDnaBlock) void k_dna_dirichlet(DnaDev m, const DnaSide *sides)
{
    const DnaSide S = sides[blockIdx.x];
    if (!m.caseActive[S.caseId]) return;
    for (int i = S.begin + threadIdx.x; i < S.end; i += kDnaBlock) {
        double nbr = 0.0;
        for (int p = m.dOff[i]; p < m.dOff[i + 1]; ++p) nbr += m.dW[p] * m.nT[m.dIdx[p]];
        const double cur = nbr + nbr * (S.K - 1.0);
        m.dPrevRes[i] = m.dRes[i];
        m.dRes[i] = cur - m.dPrev[i];
        m.dT[i] = cur;
    }
    __syncthreads();
    if (S.accType == 2) {
        iqnils(S, m.dT, m.dPrev, m.dRes, m.dHist, m.dCols, m.dCorr, S.id);
        return;
    }
    const double w = relaxFactor(S, m.dRes, m.dPrevRes, m.dOmega, m.dCorr, S.id);
    finishRelax(S, w, m.dT, m.dPrev, m.dRes, m.dOmega, m.dCorr, S.id);
}

// Neumann side: gradient = (-(gathered Dirichlet flux) + 0) / k, relaxed
__global__ __launch_bounds__(kDnaBlock) void k_dna_neumann(DnaDev m, const DnaSide *sides)
{
    const DnaSide S = sides[blockIdx.x];
    if (!m.caseActive[S.caseId]) return;
    for (int j = S.begin + threadIdx.x; j < S.end; j += kDnaBlock) {
        double fl = 0.0;
        for (int p = m.nOff[j]; p < m.nOff[j + 1]; ++p) fl += m.nW[p] * dirFlux(m, m.nIdx[p]);
        fl = fl * -1.0;
        fl = fl + 0.0;
        const double cur = fl / m.nKb[j];
        m.nPrevRes[j] = m.nRes[j];
        m.nRes[j] = cur - m.nPrev[j];
        m.nG[j] = cur;
    }
    __syncthreads();
    if (S.accType == 2) {
        iqnils(S, m.nG, m.nPrev, m.nRes, m.nHist, m.nCols, m.nCorr, S.id);
        return;
    }
    const double w = relaxFactor(S, m.nRes, m.nPrevRes, m.nOmega, m.nCorr, S.id);
    finishRelax(S, w, m.nG, m.nPrev, m.nRes, m.nOmega, m.nCorr, S.id);
}

__global__ __launch_bounds__(kDnaBlock) void k_dna_source(DnaDev m, const DnaCell *cells, int nCells,
                                                          const double *coeff, const int32_t *val, double *b)
{
    const int c = blockIdx.x * kDnaBlock + threadIdx.x;
    if (c >= nCells) return;
    const DnaCell C = cells[c];
    if (!m.caseActive[C.caseId]) return;
    double s = m.base ? m.base[C.pos] : C.base;
    for (int t = C.tBegin; t < C.tEnd; ++t) {
        const int v = val[t];
        s = s + coeff[t] * (v >= 0 ? m.dT[v] : m.nG[-v - 1]);
    }
    b[C.pos] = s;
}

// Neumann patch values after the region's solve: T_P + g / deltaCoeffs
__global__ __launch_bounds__(kDnaBlock) void k_dna_patch_values(DnaDev m, const DnaSide *sides)
{
    const DnaSide S = sides[blockIdx.x];
    if (!m.caseActive[S.caseId]) return;
    for (int j = S.begin + threadIdx.x; j < S.end; j += kDnaBlock) m.nT[j] = m.x[m.nPos[j]] + m.nG[j] / m.nDc[j];
}

__device__ inline double normRes(const double (&s)[3])
{
    const double n = fmax(fmin(sqrt(s[1]), sqrt(s[2])), kSmallFoam);
    return sqrt(s[0]) / n;
}

// jump residual on the Dirichlet side, flux residual on the Neumann side
// (genericRegionCoupledJump / genericRegionCoupledFlux normResidual)
__global__ __launch_bounds__(kDnaBlock) void k_dna_residuals(DnaDev m, const DnaSide *dSides, const DnaSide *nSides,
                                                             const DnaPair *pairs, double *out)
{
    const DnaPair P = pairs[blockIdx.x];
    if (!m.caseActive[P.caseId]) return;
    const DnaSide D = dSides[P.dir], N = nSides[P.neu];
    DD a[3] = {DD{0.0, 0.0}, DD{0.0, 0.0}, DD{0.0, 0.0}};
    for (int i = D.begin + threadIdx.x; i < D.end; i += kDnaBlock) {
        double nbr = 0.0;
        for (int p = m.dOff[i]; p < m.dOff[i + 1]; ++p) nbr += m.dW[p] * m.nT[m.dIdx[p]];
        const double jump = nbr * (D.K - 1.0);
        const double r = (m.dT[i] - nbr) - jump;
        const double o = nbr + jump;
        a[0] = ddAdd(a[0], DD{r * r, 0.0});
        a[1] = ddAdd(a[1], DD{m.dT[i] * m.dT[i], 0.0});
        a[2] = ddAdd(a[2], DD{o * o, 0.0});
    }
    double sj[3];
    blockSum<3>(a, sj);
    DD f[3] = {DD{0.0, 0.0}, DD{0.0, 0.0}, DD{0.0, 0.0}};
    for (int j = N.begin + threadIdx.x; j < N.end; j += kDnaBlock) {
        const double fown = m.nG[j] * m.nKb[j];
        double fnbr = 0.0;
        for (int p = m.nOff[j]; p < m.nOff[j + 1]; ++p) fnbr += m.nW[p] * dirFlux(m, m.nIdx[p]);
        const double r = (fown + fnbr) - 0.0;
        const double o = -1.0 * fnbr + 0.0;
        f[0] = ddAdd(f[0], DD{r * r, 0.0});
        f[1] = ddAdd(f[1], DD{fown * fown, 0.0});
        f[2] = ddAdd(f[2], DD{o * o, 0.0});
    }
    double sf[3];
    blockSum<3>(f, sf);
    if (threadIdx.x == 0) {
        out[2 * blockIdx.x] = normRes(sj);
        out[2 * blockIdx.x + 1] = normRes(sf);
    }
}

}  // namespace

hipError_t launchDnaDirichlet(const DnaDev &m, const DnaSide *sides, int nSides, hipStream_t st)
{
    if (nSides > 0) k_dna_dirichlet<<<nSides, kDnaBlock, 0, st>>>(m, sides);
    return hipGetLastError();
}

hipError_t launchDnaNeumann(const DnaDev &m, const DnaSide *sides, int nSides, hipStream_t st)
{
    if (nSides > 0) k_dna_neumann<<<nSides, kDnaBlock, 0, st>>>(m, sides);
    return hipGetLastError();
}

hipError_t launchDnaSource(const DnaDev &m, const DnaCell *cells, int nCells, const double *coeff,
                           const int32_t *val, double *b, hipStream_t st)
{
    if (nCells > 0)
        k_dna_source<<<(nCells + kDnaBlock - 1) / kDnaBlock, kDnaBlock, 0, st>>>(m, cells, nCells, coeff, val, b);
    return hipGetLastError();
}

hipError_t launchDnaPatchValues(const DnaDev &m, const DnaSide *sides, int nSides, hipStream_t st)
{
    if (nSides > 0) k_dna_patch_values<<<nSides, kDnaBlock, 0, st>>>(m, sides);
    return hipGetLastError();
}

hipError_t launchDnaResiduals(const DnaDev &m, const DnaSide *dSides, const DnaSide *nSides, const DnaPair *pairs,
                              int nPairs, double *out, hipStream_t st)
{
    if (nPairs > 0) k_dna_residuals<<<nPairs, kDnaBlock, 0, st>>>(m, dSides, nSides, pairs, out);
    return hipGetLastError();
}

}  // namespace mrfg
