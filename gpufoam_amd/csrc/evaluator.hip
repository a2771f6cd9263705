// The evaluator: device copy of the mechanism, pre-allocated buffers and
// the gfc_solve entry point.  Replaces FoamGpu::GpuKernelEvaluator
// (gpu_chemistry/src/gpuKernelEvaluator/gpuKernelEvaluator/gpuKernelEvaluator.cu:14-97)
// and its scratch resource (gpuMemoryResource.cu:7-51): everything is
// allocated once in gfc_evaluator_create, so a device-pointer solve
// allocates nothing and only enqueues kernels on the caller's stream.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <type_traits>
#include <cstring>
#include <string>
#include <vector>

#include "evaluator.hpp"
#include "mechanism.hpp"

namespace gfc {

__global__ void lane_solve_kernel(SolveArgs A, double* scratch, long per);
long lane_scratch_per_cell(int nS);

namespace {

__global__ void init_min_kernel(unsigned long long* bits) {
    *bits = static_cast<unsigned long long>(__double_as_longlong(HUGE_VAL));
}

__global__ void finish_min_kernel(const unsigned long long* bits, double* out) {
    *out = __longlong_as_double(static_cast<long long>(*bits));
}

bool hipFail(hipError_t e, const char* what) {
    if (e == hipSuccess) return false;
    setLastError(std::string(what) + ": " + hipGetErrorString(e));
    return true;
}

template <class T>
void put(std::vector<char>& blob, size_t& off, const T* src, size_t n) {
    off = (off + 255) & ~size_t(255);
    const size_t bytes = n * sizeof(T);
    if (blob.size() < off + bytes) blob.resize(off + bytes);
    if (n) std::memcpy(blob.data() + off, src, bytes);
    off += bytes;
}

} // namespace

// Uploads the mechanism tables with the reactions in the order `perm`
// (perm[r] = reference index of the reaction stored at r).
static int upload_tables(const gfc_mech_view& v, const std::vector<int>& perm,
                         DeviceMech& d, void*& block, bool blockTables = false) {
    const int nS = v.n_species, nR = v.n_reactions;
    std::vector<double> R(nS);
    for (int i = 0; i < nS; ++i) R[i] = kRR / v.W[i];
    // reaction arrays gathered in `perm` order
    auto gather = [&](auto* src, int width) {
        using T = std::remove_const_t<std::remove_pointer_t<decltype(src)>>;
        std::vector<T> out((size_t)nR * width);
        for (int r = 0; r < nR; ++r)
            for (int q = 0; q < width; ++q)
                out[(size_t)r * width + q] = src[(size_t)perm[r] * width + q];
        return out;
    };
    const auto type = gather(v.type, 1), irr = gather(v.irreversible, 1);
    const auto k = gather(v.k, 3), kInf = gather(v.kInf, 3), troe = gather(v.troe, 4);
    const auto eff = gather(v.eff, nS);
    const auto nl = gather(v.n_lhs, 1), nr = gather(v.n_rhs, 1);
    const auto li = gather(v.lhs_index, kWidth), lei = gather(v.lhs_exp_int, kWidth);
    const auto ls = gather(v.lhs_stoich, kWidth), le = gather(v.lhs_exp, kWidth);
    const auto ri = gather(v.rhs_index, kWidth), rei = gather(v.rhs_exp_int, kWidth);
    const auto rs = gather(v.rhs_stoich, kWidth), re = gather(v.rhs_exp, kWidth);
    const auto rY = gather(v.rY, 1), rW = gather(v.rW, 1), rTc = gather(v.rTcommon, 1);
    const auto rhigh = gather(v.rhigh, 7), rlow = gather(v.rlow, 7);
    const auto extra = gather(v.extra, GFC_EXTRA_PARAMS);
    int extended = 0;
    for (int r = 0; r < nR; ++r) extended |= type[r] > GFC_RATE_TROE;
    const int nPass = (nR + 63) / 64;
    std::vector<double> effP((size_t)nPass * nS * 64, 0.0);
    for (int r = 0; r < nR; ++r) {
        const int pass = r / 64, lane = r % 64;
        for (int i = 0; i < nS; ++i)
            effP[((size_t)pass * nS + i) * 64 + lane] = eff[(size_t)r * nS + i];
    }
    // rate constants in log form and de-duplicated efficiency vectors
    std::vector<double> lnA(2 * (size_t)nR), sgnA(2 * (size_t)nR);
    for (int r = 0; r < nR; ++r) {
        for (int q = 0; q < 2; ++q) {
            const double A = (q == 0 ? k : kInf)[3 * (size_t)r];
            lnA[2 * r + q] = std::log(std::fabs(A));
            sgnA[2 * r + q] = A < 0 ? -1.0 : 1.0;
        }
    }
    std::vector<std::vector<double>> sets;
    std::vector<int> effSet(nR, -1);
    for (int r = 0; r < nR; ++r) {
        if (!uses_third_body(type[r])) continue;
        std::vector<double> row(eff.begin() + (size_t)r * nS, eff.begin() + (size_t)(r + 1) * nS);
        int s = -1;
        for (size_t q = 0; q < sets.size(); ++q)
            if (sets[q] == row) s = static_cast<int>(q);
        if (s < 0) {
            s = static_cast<int>(sets.size());
            sets.push_back(row);
        }
        effSet[r] = s;
    }
    std::vector<double> effSets(std::max<size_t>(1, sets.size()) * 64, 0.0);
    for (size_t q = 0; q < sets.size(); ++q)
        for (int i = 0; i < nS && i < 64; ++i) effSets[q * 64 + i] = sets[q][i];
    std::vector<double> effSetsT((size_t)64 * kSetsT, 0.0);
    for (size_t q = 0; q < sets.size() && q < (size_t)kSetsT; ++q)
        for (int i = 0; i < nS && i < 64; ++i) effSetsT[(size_t)i * kSetsT + q] = sets[q][i];
    // wavefront-kernel records (only usable when no side is wider than
    // kTeamWidth)
    int teamWidth = 0;
    for (int r = 0; r < nR; ++r) teamWidth = std::max({teamWidth, nl[r], nr[r]});
    std::vector<double> recD((size_t)nPass * kRecD * 64, 0.0);
    std::vector<int> recI((size_t)nPass * kRecI * 64, 0);
    for (int r = 0; r < nR; ++r) {
        const size_t pass = r / 64, lane = r % 64;
        auto D = [&](int f) -> double& { return recD[(pass * kRecD + f) * 64 + lane]; };
        auto I = [&](int f) -> int& { return recI[(pass * kRecI + f) * 64 + lane]; };
        D(rdLnA) = lnA[2 * r];
        D(rdBeta) = k[3 * r + 1];
        D(rdTa) = k[3 * r + 2];
        D(rdSgnA) = sgnA[2 * r];
        D(rdLnAi) = lnA[2 * r + 1];
        D(rdBetai) = kInf[3 * r + 1];
        D(rdTai) = kInf[3 * r + 2];
        D(rdSgnAi) = sgnA[2 * r + 1];
        for (int q = 0; q < 4; ++q) D(rdAlpha + q) = troe[4 * r + q];
        double dn = 0.0;
        for (int q = 0; q < nl[r]; ++q) dn -= ls[kWidth * r + q];
        for (int q = 0; q < nr[r]; ++q) dn += rs[kWidth * r + q];
        D(rdDn) = dn;
        D(rdTcommon) = rTc[r];
        I(riType) = type[r];
        I(riIrr) = irr[r];
        I(riSet) = effSet[r] < 0 ? 0 : effSet[r];
        I(riNl) = nl[r];
        I(riNr) = nr[r];
        for (int q = 0; q < kTeamWidth; ++q) {
            const bool okl = q < nl[r], okr = q < nr[r];
            D(rdLs + q) = okl ? ls[kWidth * r + q] : 0.0;
            D(rdRs + q) = okr ? rs[kWidth * r + q] : 0.0;
            D(rdLe + q) = okl ? le[kWidth * r + q] : 0.0;
            D(rdRe + q) = okr ? re[kWidth * r + q] : 0.0;
            I(riLi + q) = okl ? li[kWidth * r + q] : 0;
            I(riRi + q) = okr ? ri[kWidth * r + q] : 0;
            I(riLei + q) = okl ? lei[kWidth * r + q] : 0;
            I(riRei + q) = okr ? rei[kWidth * r + q] : 0;
        }
    }
    // gather lists of the workgroup-per-cell kernel
    std::vector<int> spPtr(1, 0), spR, elPtr(1, 0), elRS, tbPtr(1, 0), tbR, tbSet;
    std::vector<double> spNu, elNu, tbNu, effB, effT;
    if (blockTables) {
        // net coefficient of specie i in reaction r
        auto nu = [&](int r, int i) {
            double v = 0.0;
            for (int q = 0; q < nl[r]; ++q)
                if (li[kWidth * r + q] == i) v -= ls[kWidth * r + q];
            for (int q = 0; q < nr[r]; ++q)
                if (ri[kWidth * r + q] == i) v += rs[kWidth * r + q];
            return v;
        };
        std::vector<std::vector<std::pair<int, double>>> rows(nS);
        for (int r = 0; r < nR; ++r) {
            std::vector<int> sp;
            for (int q = 0; q < nl[r]; ++q) sp.push_back(li[kWidth * r + q]);
            for (int q = 0; q < nr[r]; ++q) sp.push_back(ri[kWidth * r + q]);
            std::sort(sp.begin(), sp.end());
            sp.erase(std::unique(sp.begin(), sp.end()), sp.end());
            for (int i : sp) {
                const double c = nu(r, i);
                if (c != 0.0) rows[i].push_back({r, c});
            }
        }
        for (int i = 0; i < nS; ++i) {
            for (auto& e : rows[i]) {
                spR.push_back(e.first);
                spNu.push_back(e.second);
                if (type[e.first] == GFC_RATE_THIRD_BODY) {
                    tbR.push_back(e.first);
                    tbSet.push_back(effSet[e.first]);
                    tbNu.push_back(e.second);
                }
            }
            spPtr.push_back((int)spR.size());
            tbPtr.push_back((int)tbR.size());
        }
        // entry lists: for each (i, j), the reactions of row i in which
        // specie j is a reactant (slot q) or product (slot 3+q)
        std::vector<std::vector<std::pair<int, double>>> el((size_t)nS * nS);
        for (int i = 0; i < nS; ++i)
            for (auto& e : rows[i]) {
                const int r = e.first;
                for (int q = 0; q < nl[r]; ++q)
                    el[(size_t)i * nS + li[kWidth * r + q]].push_back({r * 8 + q, e.second});
                for (int q = 0; q < nr[r]; ++q)
                    el[(size_t)i * nS + ri[kWidth * r + q]].push_back({r * 8 + 3 + q, e.second});
            }
        elPtr.reserve((size_t)nS * nS + 1);
        for (size_t k = 0; k < el.size(); ++k) {
            for (auto& e : el[k]) {
                elRS.push_back(e.first);
                elNu.push_back(e.second);
            }
            elPtr.push_back((int)elRS.size());
        }
        const size_t nSet = sets.size();
        effB.assign(std::max<size_t>(1, nSet) * nS, 0.0);
        effT.assign((size_t)nS * std::max<size_t>(1, nSet), 0.0);
        for (size_t q = 0; q < nSet; ++q)
            for (int i = 0; i < nS; ++i) {
                effB[q * nS + i] = sets[q][i];
                effT[(size_t)i * nSet + q] = sets[q][i];
            }
    }
    std::vector<char> blob;
    size_t off = 0;
    std::vector<size_t> offs;
    auto add = [&](const auto* p, size_t n) {
        put(blob, off, p, n);
        offs.push_back(off - n * sizeof(*p));
    };
    add(v.W, nS);
    add(R.data(), nS);
    add(v.sTcommon, nS);
    add(v.high, 7 * nS);
    add(v.low, 7 * nS);
    add(type.data(), nR);
    add(irr.data(), nR);
    add(k.data(), 3 * nR);
    add(kInf.data(), 3 * nR);
    add(troe.data(), 4 * nR);
    add(eff.data(), (size_t)nS * nR);
    add(nl.data(), nR);
    add(nr.data(), nR);
    add(li.data(), kWidth * nR);
    add(ls.data(), kWidth * nR);
    add(le.data(), kWidth * nR);
    add(lei.data(), kWidth * nR);
    add(ri.data(), kWidth * nR);
    add(rs.data(), kWidth * nR);
    add(re.data(), kWidth * nR);
    add(rei.data(), kWidth * nR);
    add(rY.data(), nR);
    add(rW.data(), nR);
    add(rTc.data(), nR);
    add(rhigh.data(), 7 * nR);
    add(rlow.data(), 7 * nR);
    add(effP.data(), effP.size());
    add(lnA.data(), lnA.size());
    add(sgnA.data(), sgnA.size());
    add(effSets.data(), effSets.size());
    add(effSetsT.data(), effSetsT.size());
    add(effSet.data(), effSet.size());
    add(perm.data(), perm.size());
    add(recD.data(), recD.size());
    add(recI.data(), recI.size());
    add(spPtr.data(), spPtr.size());
    add(spR.data(), spR.size());
    add(spNu.data(), spNu.size());
    add(elPtr.data(), elPtr.size());
    add(elRS.data(), elRS.size());
    add(elNu.data(), elNu.size());
    add(tbPtr.data(), tbPtr.size());
    add(tbR.data(), tbR.size());
    add(tbSet.data(), tbSet.size());
    add(tbNu.data(), tbNu.size());
    add(effB.data(), effB.size());
    add(effT.data(), effT.size());
    add(extra.data(), extra.size());
    if (hipFail(hipMalloc(&block, blob.size()), "hipMalloc(mechanism)"))
        return GFC_ERR_HIP;
    if (hipFail(hipMemcpy(block, blob.data(), blob.size(), hipMemcpyHostToDevice),
                "hipMemcpy(mechanism)"))
        return GFC_ERR_HIP;
    char* base = static_cast<char*>(block);
    int q = 0;
    auto at = [&](auto*& dst) {
        dst = reinterpret_cast<std::remove_reference_t<decltype(dst)>>(base + offs[q++]);
    };
    d.nS = nS;
    d.nR = nR;
    d.Tlow = v.Tlow;
    d.Thigh = v.Thigh;
    at(d.W); at(d.R); at(d.Tcommon); at(d.high); at(d.low);
    at(d.type); at(d.irreversible); at(d.k); at(d.kInf); at(d.troe); at(d.eff);
    at(d.nl); at(d.nr);
    at(d.li); at(d.ls); at(d.le); at(d.lei);
    at(d.ri); at(d.rs); at(d.re); at(d.rei);
    at(d.rY); at(d.rW); at(d.rTcommon); at(d.rhigh); at(d.rlow);
    at(d.effP);
    at(d.lnA);
    at(d.sgnA);
    at(d.effSets);
    at(d.effSetsT);
    at(d.effSet);
    at(d.rxnRef);
    at(d.recD);
    at(d.recI);
    at(d.blk.spPtr); at(d.blk.spR); at(d.blk.spNu);
    at(d.blk.elPtr); at(d.blk.elRS); at(d.blk.elNu);
    at(d.blk.tbPtr); at(d.blk.tbR); at(d.blk.tbSet); at(d.blk.tbNu);
    at(d.blk.effB); at(d.blk.effT);
    at(d.extra);
    d.extended = extended;
    d.teamWidth = teamWidth;
    d.nPass = nPass;
    d.nSets = static_cast<int>(sets.size());
    return GFC_OK;
}

SolveArgs solve_args(const gfc_evaluator* ev, double deltaT, double deltaTChemMax,
                     int nCells) {
    SolveArgs A{};
    A.m = ev->dm;
    A.tab = tableau(ev->ode.solver);
    A.absTol = ev->ode.absTol;
    A.relTol = ev->ode.relTol;
    A.maxSteps = ev->ode.maxSteps;
    A.safeScale = ev->ode.safeScale;
    A.alphaInc = ev->ode.alphaInc;
    A.alphaDec = ev->ode.alphaDec;
    A.minScale = ev->ode.minScale;
    A.maxScale = ev->ode.maxScale;
    A.incThreshold = std::pow(ev->ode.maxScale / ev->ode.safeScale, -1.0 / ev->ode.alphaInc);
    A.Treact = ev->Treact;
    A.deltaT = deltaT;
    A.deltaTChemMax = deltaTChemMax;
    A.nCells = nCells;
    A.qChunk = 1;
    A.qBulk = 0;
    return A;
}

// Mechanisms the workgroup-per-cell kernel can run (its distinct
// third-body sets are checked again at launch).
bool block_eligible(const gfc_mech_view& v) {
    const int n = v.n_species + 1;
    if (n <= 64 || n > kBlockMaxN || v.n_reactions > kBlockMaxR) return false;
    // the per-reaction derivatives (6 per reaction) share LDS with the
    // diagonal factor blocks during the Jacobian assembly
    const long nRp = ((long)v.n_reactions + 63) / 64 * 64;
    if (6 * nRp > block_lds_doubles(n)) return false;
    int width = 0, nonArr = 0;
    for (int r = 0; r < v.n_reactions; ++r) {
        width = std::max({width, v.n_lhs[r], v.n_rhs[r]});
        nonArr += v.type[r] != GFC_RATE_ARRHENIUS;
        if (v.type[r] > GFC_RATE_TROE) return false; // gpuFoam's four types only
    }
    return width <= kTeamWidth && nonArr <= kBlockMaxSets;
}

// Uploads both table orders from the evaluator's copy of the mechanism
// (replacing earlier uploads).
int evaluator_upload(gfc_evaluator* ev) {
    if (ev->mechBlock) (void)hipFree(ev->mechBlock);
    if (ev->mechBlockTeam) (void)hipFree(ev->mechBlockTeam);
    ev->mechBlock = ev->mechBlockTeam = nullptr;
    const gfc_mech_view v = ev->mech.view();
    const int nR = v.n_reactions;
    // reference order for the per-lane kernel
    std::vector<int> ident(nR);
    for (int r = 0; r < nR; ++r) ident[r] = r;
    int st = upload_tables(v, ident, ev->dm, ev->mechBlock);
    if (st != GFC_OK) return st;
    // wavefront kernel: reactions grouped by rate type so that a pass of 64
    // reactions follows one code path
    std::vector<int> perm = ident;
    auto rank = [&](int r) {
        return v.type[r] * 2 + (v.type[r] == GFC_RATE_ARRHENIUS ? v.irreversible[r] : 0);
    };
    std::stable_sort(perm.begin(), perm.end(),
                     [&](int a, int b) { return rank(a) < rank(b); });
    st = upload_tables(v, perm, ev->dmTeam, ev->mechBlockTeam, block_eligible(v));
    if (st != GFC_OK) return st;
    ev->dm.bounds = ev->dmTeam.bounds = ev->boundsWord;
    return GFC_OK;
}

int bounds_report(gfc_evaluator* ev, hipStream_t s) {
    if (!kBoundsChecked || !ev->boundsWord) return GFC_OK;
    unsigned w = 0;
    if (hipFail(hipMemcpyAsync(&w, ev->boundsWord, sizeof w, hipMemcpyDeviceToHost, s),
                "hipMemcpyAsync(bounds)") ||
        hipFail(hipStreamSynchronize(s), "hipStreamSynchronize"))
        return GFC_ERR_HIP;
    if (w == 0) return GFC_OK;
    const unsigned z = 0;
    (void)hipMemcpy(ev->boundsWord, &z, sizeof z, hipMemcpyHostToDevice);
    char msg[160];
    snprintf(msg, sizeof msg,
             "bounds check failed (mask 0x%x: 1 specie, 2 efficiency set, 4 cell, "
             "8 recorded step, 16 checkpoint slot)", w);
    setLastError(msg);
    return GFC_ERR_BOUNDS;
}

} // namespace gfc

using namespace gfc;

extern "C" {

int gfc_evaluator_create(const gfc_mechanism* mech, const gfc_ode_inputs* ode,
                         double Treact, int maxCells, gfc_evaluator** out) {
    if (!mech || !ode || !out || maxCells <= 0) {
        setLastError("gfc_evaluator_create: bad argument");
        return GFC_ERR_INVALID_ARGUMENT;
    }
    if (ode->solver != 12 && ode->solver != 23 && ode->solver != 34) {
        setLastError("Invalid ODE Solver type");
        return GFC_ERR_SOLVER_NAME;
    }
    *out = nullptr;
    auto* ev = new gfc_evaluator();
    ev->nS = mech->nS;
    ev->nR = mech->nR;
    ev->maxCells = maxCells;
    ev->ode = *ode;
    ev->Treact = Treact;
    hipGetDevice(&ev->device);
    // the caller-free entry point's device array, resolved on this device
    int sst = noalloc_init_device(ev->device);
    if (sst != GFC_OK) {
        delete ev;
        return sst;
    }
    ev->mech = *mech;
    if (kBoundsChecked) {
        if (hipFail(hipMalloc(&ev->boundsWord, sizeof(unsigned)), "hipMalloc(bounds)") ||
            hipFail(hipMemset(ev->boundsWord, 0, sizeof(unsigned)), "hipMemset(bounds)")) {
            gfc_evaluator_destroy(ev);
            return GFC_ERR_HIP;
        }
    }
    int st = evaluator_upload(ev);
    if (st != GFC_OK) {
        gfc_evaluator_destroy(ev);
        return st;
    }
    // plain kernel scratch: one slot per cell, rounded to whole wavefronts
    ev->laneStride = ((long)maxCells + 63) / 64 * 64;
    ev->lanePer = lane_scratch_per_cell(ev->nS);
    ev->laneScratchBytes = sizeof(double) * ev->lanePer * ev->laneStride;
    const int nE = ev->nS + 2;
    const size_t ld = ev->laneStride;
    if (hipFail(hipMalloc(&ev->laneScratch, ev->laneScratchBytes), "hipMalloc(scratch)") ||
        hipFail(hipMalloc(&ev->minBits, 256), "hipMalloc(min)") ||
        hipFail(hipMalloc(&ev->sY0, sizeof(double) * nE * ld), "hipMalloc(staging)") ||
        hipFail(hipMalloc(&ev->sYnew, sizeof(double) * nE * ld), "hipMalloc(staging)") ||
        hipFail(hipMalloc(&ev->sRR, sizeof(double) * ev->nS * ld), "hipMalloc(staging)") ||
        hipFail(hipMalloc(&ev->sRho, sizeof(double) * ld), "hipMalloc(staging)") ||
        hipFail(hipMalloc(&ev->sDt, sizeof(double) * ld), "hipMalloc(staging)") ||
        hipFail(hipMalloc(&ev->sDtMin, 256), "hipMalloc(staging)") ||
        hipFail(hipMalloc(&ev->sStats, sizeof(int) * 3 * ld), "hipMalloc(staging)")) {
        gfc_evaluator_destroy(ev);
        return GFC_ERR_HIP;
    }
    st = batched_init(ev);
    if (st != GFC_OK) {
        gfc_evaluator_destroy(ev);
        return st;
    }
    *out = ev;
    return GFC_OK;
}

void gfc_evaluator_destroy(gfc_evaluator* ev) {
    if (!ev) return;
    batched_free(ev);
    adjoint_free(ev);
    void* ptrs[] = {ev->mechBlock, ev->mechBlockTeam, ev->laneScratch, ev->minBits, ev->sY0,
                    ev->sYnew, ev->sRR, ev->sRho, ev->sDt, ev->sDtMin,
                    ev->sStats, ev->boundsWord};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    delete ev;
}

int gfc_solve(gfc_evaluator* ev, double deltaT, double deltaTChemMax,
              int nCells, int ldY, const double* rho, const double* Y0,
              double* Ynew, double* RR, double* deltaTChem, double* deltaTMin,
              int* cellStats, void* stream, unsigned flags) {
    // with no cells the field pointers may be null (zero-size buffers)
    if (!ev || nCells < 0 || ldY < nCells || !(deltaT > 0) ||
        (nCells > 0 && (!Y0 || !deltaTChem || (RR && !rho)))) {
        setLastError("gfc_solve: bad argument");
        return GFC_ERR_INVALID_ARGUMENT;
    }
    if (nCells > ev->maxCells) {
        setLastError("gfc_solve: nCells exceeds the evaluator's maxCells");
        return GFC_ERR_CAPACITY;
    }
    if (nCells == 0) {
        if (deltaTMin) {
            if (flags & GFC_HOST_POINTERS) *deltaTMin = HUGE_VAL;
            else {
                hipStream_t s = static_cast<hipStream_t>(stream);
                hipLaunchKernelGGL(init_min_kernel, dim3(1), dim3(1), 0, s, ev->minBits);
                hipLaunchKernelGGL(finish_min_kernel, dim3(1), dim3(1), 0, s,
                                   ev->minBits, deltaTMin);
            }
        }
        return GFC_OK;
    }
    hipStream_t s = static_cast<hipStream_t>(stream);
    const int nS = ev->nS, nE = nS + 2;
    const bool host = flags & GFC_HOST_POINTERS;

    SolveArgs A = solve_args(ev, deltaT, deltaTChemMax, nCells);

    if (host) {
        // stage host buffers through the evaluator's device buffers
        const long ld = nCells;
        for (int k = 0; k < nE; ++k)
            if (hipFail(hipMemcpyAsync(ev->sY0 + k * ld, Y0 + (long)k * ldY,
                                       sizeof(double) * nCells,
                                       hipMemcpyHostToDevice, s), "hipMemcpyAsync"))
                return GFC_ERR_HIP;
        if (hipFail(hipMemcpyAsync(ev->sDt, deltaTChem, sizeof(double) * nCells,
                                   hipMemcpyHostToDevice, s), "hipMemcpyAsync"))
            return GFC_ERR_HIP;
        if (RR && hipFail(hipMemcpyAsync(ev->sRho, rho, sizeof(double) * nCells,
                                         hipMemcpyHostToDevice, s), "hipMemcpyAsync"))
            return GFC_ERR_HIP;
        A.ldY = nCells;
        A.rho = RR ? ev->sRho : nullptr;
        A.Y0 = ev->sY0;
        A.Ynew = Ynew ? ev->sYnew : nullptr;
        A.RR = RR ? ev->sRR : nullptr;
        A.dtChem = ev->sDt;
        A.stats = cellStats ? ev->sStats : nullptr;
    } else {
        A.ldY = ldY;
        A.rho = rho;
        A.Y0 = Y0;
        A.Ynew = Ynew;
        A.RR = RR;
        A.dtChem = deltaTChem;
        A.stats = cellStats;
    }
    A.dtMinBits = deltaTMin ? ev->minBits : nullptr;
    if (deltaTMin)
        hipLaunchKernelGGL(init_min_kernel, dim3(1), dim3(1), 0, s, ev->minBits);

    int st = GFC_OK;
    if (flags & GFC_BATCHED) {
        st = batched_launch(ev, A, s, (flags & GFC_GRAPH) != 0);
        if (st != GFC_OK) return st;
    } else {
        st = plain_launch(ev, A, s, (flags & GFC_LANE_PER_CELL) != 0);
        if (st != GFC_OK) return st;
    }
    if (hipFail(hipGetLastError(), "kernel launch")) return GFC_ERR_HIP;
    if (deltaTMin)
        hipLaunchKernelGGL(finish_min_kernel, dim3(1), dim3(1), 0, s, ev->minBits,
                           host ? ev->sDtMin : deltaTMin);

    if (host) {
        const long ld = nCells;
        if (Ynew)
            for (int k = 0; k < nE; ++k)
                hipMemcpyAsync(Ynew + (long)k * ldY, ev->sYnew + k * ld,
                               sizeof(double) * nCells, hipMemcpyDeviceToHost, s);
        if (RR)
            for (int k = 0; k < nS; ++k)
                hipMemcpyAsync(RR + (long)k * ldY, ev->sRR + k * ld,
                               sizeof(double) * nCells, hipMemcpyDeviceToHost, s);
        hipMemcpyAsync(deltaTChem, ev->sDt, sizeof(double) * nCells,
                       hipMemcpyDeviceToHost, s);
        if (deltaTMin)
            hipMemcpyAsync(deltaTMin, ev->sDtMin, sizeof(double), hipMemcpyDeviceToHost, s);
        if (cellStats)
            for (int k = 0; k < 3; ++k)
                hipMemcpyAsync(cellStats + (long)k * ldY, ev->sStats + k * ld,
                               sizeof(int) * nCells, hipMemcpyDeviceToHost, s);
        if (hipFail(hipStreamSynchronize(s), "hipStreamSynchronize")) return GFC_ERR_HIP;
    }
    if (st == GFC_OK) st = bounds_report(ev, s);
    return st;
}

int gfc_bounds_checked(void) { return kBoundsChecked ? 1 : 0; }

} // extern "C"
