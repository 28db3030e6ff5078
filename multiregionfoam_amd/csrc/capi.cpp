// extern "C" entry points of libmrfg.so (declared in include/mrfg.h).
#include "../../include/mrfg.h"

#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <tuple>
#include <vector>

#include "batch.hpp"
#include "dna_run.hpp"
#include "noalloc.hpp"
#include "serve.hpp"
#include "coupled_system.hpp"
#include "dictionary.hpp"

struct mrfg_ctx {
    int device = 0;
    std::string err;
    std::vector<std::unique_ptr<mrfg::CoupledSystem>> systems;
    std::unique_ptr<mrfg::Batch> batch;
    std::unique_ptr<mrfg::DnaRun> dna;   // partitioned coupling of the committed systems
    bool committed = false;
    bool preferLevels = false;
};

struct mrfg_case {
    mrfg::CoupledSystem sys;
};

namespace {

struct StateError : std::runtime_error {
    using std::runtime_error::runtime_error;
};
struct ArgError : std::runtime_error {
    using std::runtime_error::runtime_error;
};

template <typename F>
int guard(mrfg_ctx *ctx, F &&f)
{
    try {
        f();
        if (ctx) ctx->err.clear();
        return MRFG_OK;
    } catch (const mrfg::DictError &e) {
        if (ctx) ctx->err = e.what();
        return MRFG_ERR_DICT;
    } catch (const StateError &e) {
        if (ctx) ctx->err = e.what();
        return MRFG_ERR_STATE;
    } catch (const ArgError &e) {
        if (ctx) ctx->err = e.what();
        return MRFG_ERR_ARG;
    } catch (const std::out_of_range &e) {
        if (ctx) ctx->err = std::string("index out of range: ") + e.what();
        return MRFG_ERR_ARG;
    } catch (const std::exception &e) {
        if (ctx) ctx->err = e.what();
        std::string w = e.what();
        if (w.rfind("HIP error", 0) == 0) return MRFG_ERR_DEVICE;
        return MRFG_ERR_ARG;
    } catch (...) {
        if (ctx) ctx->err = "unknown error";
        return MRFG_ERR_ARG;
    }
}

void copyErr(const std::string &m, char *err, int len)
{
    if (err && len > 0) {
        std::strncpy(err, m.c_str(), (size_t)len - 1);
        err[len - 1] = '\0';
    }
}

mrfg::CoupledSystem &systemOf(mrfg_ctx *ctx, int sys)
{
    if (sys < 0 || sys >= (int)ctx->systems.size()) throw ArgError("system index out of range");
    return *ctx->systems[sys];
}

mrfg::Batch &committedBatch(mrfg_ctx *ctx)
{
    if (!ctx->committed || !ctx->batch) throw StateError("mrfg_commit has not been called");
    return *ctx->batch;
}

void toControls(const mrfg::SolverControls &s, mrfg_controls *c)
{
    c->solver = s.solver == mrfg::SolverControls::PCG ? MRFG_PCG : MRFG_BICGSTAB;
    c->precision = s.precision == mrfg::SolverControls::Mixed ? MRFG_MIXED : MRFG_FP64;
    c->tolerance = s.tolerance;
    c->rel_tol = s.relTol;
    c->min_iter = s.minIter;
    c->max_iter = s.maxIter;
    c->inner_rel_tol = s.innerRelTol;
    c->inner_max_iter = s.innerMaxIter;
    c->max_outer = s.maxOuter;
    c->check_interval = s.checkInterval;
    c->launch_mode = s.graphs ? MRFG_LAUNCH_GRAPH : MRFG_LAUNCH_KERNELS;
    c->tuning = s.autotune ? MRFG_TUNE_AUTO : MRFG_TUNE_FIXED;
    c->inner_preconditioner = s.innerPre == mrfg::SolverControls::InnerDILU ? MRFG_INNER_DILU : MRFG_INNER_TILE_DILU;
}

mrfg::SolverControls fromControls(const mrfg_controls *c)
{
    mrfg::SolverControls s;
    if (c->solver != MRFG_BICGSTAB && c->solver != MRFG_PCG) throw ArgError("unknown solver id");
    if (c->precision != MRFG_FP64 && c->precision != MRFG_MIXED) throw ArgError("unknown precision id");
    s.solver = c->solver == MRFG_PCG ? mrfg::SolverControls::PCG : mrfg::SolverControls::BiCGStab;
    s.precision = c->precision == MRFG_MIXED ? mrfg::SolverControls::Mixed : mrfg::SolverControls::FP64;
    s.tolerance = c->tolerance;
    s.relTol = c->rel_tol;
    s.minIter = c->min_iter;
    s.maxIter = c->max_iter;
    s.innerRelTol = c->inner_rel_tol;
    s.innerMaxIter = c->inner_max_iter;
    s.maxOuter = c->max_outer;
    s.checkInterval = c->check_interval > 0 ? c->check_interval : 8;
    if (c->launch_mode != MRFG_LAUNCH_GRAPH && c->launch_mode != MRFG_LAUNCH_KERNELS) throw ArgError("unknown launch mode");
    s.graphs = c->launch_mode == MRFG_LAUNCH_GRAPH;
    if (c->tuning != MRFG_TUNE_AUTO && c->tuning != MRFG_TUNE_FIXED) throw ArgError("unknown tuning mode");
    s.autotune = c->tuning == MRFG_TUNE_AUTO;
    if (c->inner_preconditioner != MRFG_INNER_DILU && c->inner_preconditioner != MRFG_INNER_TILE_DILU)
        throw ArgError("unknown inner preconditioner id");
    s.innerPre = c->inner_preconditioner == MRFG_INNER_DILU ? mrfg::SolverControls::InnerDILU
                                                            : mrfg::SolverControls::InnerTileDILU;
    if (s.maxIter < 0 || s.minIter < 0 || s.innerMaxIter < 1 || s.maxOuter < 1) throw ArgError("bad iteration limits");
    return s;
}

mrfg::SolverControls controlsFromText(const char *text, const char *entry)
{
    if (!text || !entry) throw ArgError("null fvSolution text or entry");
    mrfg::Dict d = mrfg::parseDict(text);
    const mrfg::Dict &solvers = d.subDict("solvers");
    return mrfg::parseControls(solvers.subDict(entry));
}

void fillPerf(const std::vector<mrfg::SolvePerf> &p, mrfg_perf *out)
{
    for (size_t i = 0; i < p.size(); ++i) {
        out[i].initial_residual = p[i].initialResidual;
        out[i].final_residual = p[i].finalResidual;
        out[i].n_iterations = p[i].nIterations;
        out[i].converged = p[i].converged;
        out[i].singular = p[i].singular;
        out[i].n_outer = p[i].nOuter;
        out[i].seconds = p[i].seconds;
        out[i].bytes_moved = p[i].bytes;
    }
}

}  // namespace

extern "C" {

const char *mrfg_version(void) { return "multiregionfoam_amd 0.1 (gfx950)"; }

int mrfg_create(mrfg_ctx **ctx, int device)
{
    if (!ctx) return MRFG_ERR_ARG;
    *ctx = nullptr;
    auto c = std::make_unique<mrfg_ctx>();
    c->device = device;
    int rc = guard(c.get(), [&] {
        int n = 0;
        if (hipGetDeviceCount(&n) != hipSuccess || n == 0) throw std::runtime_error("HIP error: no HIP device");
        if (device < 0 || device >= n) throw ArgError("device index out of range");
        c->batch = std::make_unique<mrfg::Batch>(device);
        (void)mrfg::fluxScratch(device);   // the allocation-free flux check's scratch, once per device
    });
    if (rc != MRFG_OK) return rc;
    *ctx = c.release();
    return MRFG_OK;
}

void mrfg_destroy(mrfg_ctx *ctx) { delete ctx; }

const char *mrfg_last_error(const mrfg_ctx *ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int mrfg_begin_system(mrfg_ctx *ctx, int n_regions, int n_interfaces, int *sys_id)
{
    if (!ctx) return MRFG_ERR_ARG;
    return guard(ctx, [&] {
        if (n_regions < 1 || n_interfaces < 0) throw ArgError("need at least one region");
        auto s = std::make_unique<mrfg::CoupledSystem>();
        s->regions.resize(n_regions);
        s->interfaces.resize(n_interfaces);
        ctx->systems.push_back(std::move(s));
        ctx->committed = false;
        if (sys_id) *sys_id = (int)ctx->systems.size() - 1;
    });
}

int mrfg_set_region(mrfg_ctx *ctx, int sys, int region, int n_cells, int n_faces, const int32_t *l_addr,
                    const int32_t *u_addr, const double *diag, const double *lower, const double *upper,
                    const double *source, const double *psi0)
{
    if (!ctx) return MRFG_ERR_ARG;
    return guard(ctx, [&] {
        mrfg::CoupledSystem &S = systemOf(ctx, sys);
        if (region < 0 || region >= (int)S.regions.size()) throw ArgError("region index out of range");
        if (n_cells < 1 || n_faces < 0) throw ArgError("bad region sizes");
        if (!diag || !source || !psi0 || (n_faces > 0 && (!l_addr || !u_addr || !upper)))
            throw ArgError("null region array");
        mrfg::LduRegion &R = S.regions[region];
        R.name = "region" + std::to_string(region);
        R.nCells = n_cells;
        R.nFaces = n_faces;
        R.lAddr.assign(l_addr, l_addr + n_faces);
        R.uAddr.assign(u_addr, u_addr + n_faces);
        R.diag.assign(diag, diag + n_cells);
        R.upper.assign(upper, upper + n_faces);
        R.symmetric = lower == nullptr;
        if (lower) R.lower.assign(lower, lower + n_faces);
        else R.lower = R.upper;
        R.source.assign(source, source + n_cells);
        R.psi.assign(psi0, psi0 + n_cells);
        ctx->committed = false;
    });
}

int mrfg_set_region_ddt(mrfg_ctx *ctx, int sys, int region, const double *ddt_coeff, const double *diag_base,
                        const double *source_base)
{
    if (!ctx) return MRFG_ERR_ARG;
    return guard(ctx, [&] {
        mrfg::CoupledSystem &S = systemOf(ctx, sys);
        if (region < 0 || region >= (int)S.regions.size()) throw ArgError("region index out of range");
        mrfg::LduRegion &R = S.regions[region];
        if (R.nCells < 1) throw ArgError("set the region before its ddt terms");
        if (!ddt_coeff) {
            R.ddtCoeff.clear();
            R.diagBase.clear();
            R.sourceBase.clear();
        } else {
            if (!diag_base || !source_base) throw ArgError("null ddt array");
            R.ddtCoeff.assign(ddt_coeff, ddt_coeff + R.nCells);
            R.diagBase.assign(diag_base, diag_base + R.nCells);
            R.sourceBase.assign(source_base, source_base + R.nCells);
        }
        ctx->committed = false;
    });
}

int mrfg_set_interface(mrfg_ctx *ctx, int sys, int iface, int region_a, int region_b, int n_faces_a,
                       const int32_t *fc_a, const double *coeff_a, const int32_t *off_a, const int32_t *idx_a,
                       const double *w_a, int n_faces_b, const int32_t *fc_b, const double *coeff_b,
                       const int32_t *off_b, const int32_t *idx_b, const double *w_b)
{
    if (!ctx) return MRFG_ERR_ARG;
    return guard(ctx, [&] {
        mrfg::CoupledSystem &S = systemOf(ctx, sys);
        if (iface < 0 || iface >= (int)S.interfaces.size()) throw ArgError("interface index out of range");
        if (n_faces_a < 0 || n_faces_b < 0) throw ArgError("bad interface sizes");
        if (!off_a || !off_b) throw ArgError("null GGI offsets");
        mrfg::CoupledInterface &I = S.interfaces[iface];
        I.name = "interface" + std::to_string(iface);
        auto fill = [](mrfg::InterfaceSide &s, int region, int n, const int32_t *fc, const double *coeff,
                       const int32_t *off, const int32_t *idx, const double *w) {
            if (n > 0 && (!fc || !coeff)) throw ArgError("null interface array");
            s.region = region;
            s.fc.assign(fc, fc + n);
            s.coeff.assign(coeff, coeff + n);
            s.magSf.assign(n, 0.0);
            s.kOwn.clear();   // wall conductances: mrfg_set_interface_wall
            s.kNei.clear();
            s.off.assign(off, off + n + 1);
            int np = s.off.back();
            if (np < 0 || s.off.front() != 0) throw ArgError("bad GGI offsets");
            for (int i = 0; i < n; ++i)
                if (s.off[i + 1] < s.off[i]) throw ArgError("GGI offsets must be non-decreasing");
            if (np > 0 && (!idx || !w)) throw ArgError("null GGI pair array");
            s.idx.assign(idx, idx + np);
            s.w.assign(w, w + np);
        };
        fill(I.a, region_a, n_faces_a, fc_a, coeff_a, off_a, idx_a, w_a);
        fill(I.b, region_b, n_faces_b, fc_b, coeff_b, off_b, idx_b, w_b);
        ctx->committed = false;
    });
}

int mrfg_set_schedule(mrfg_ctx *ctx, int schedule)
{
    if (!ctx) return MRFG_ERR_ARG;
    return guard(ctx, [&] {
        if (schedule != MRFG_SCHEDULE_AUTO && schedule != MRFG_SCHEDULE_LEVELS) throw ArgError("unknown schedule");
        ctx->batch->setPreferLevels(schedule == MRFG_SCHEDULE_LEVELS);
        ctx->preferLevels = schedule == MRFG_SCHEDULE_LEVELS;
        ctx->committed = false;
    });
}

int mrfg_commit(mrfg_ctx *ctx)
{
    if (!ctx) return MRFG_ERR_ARG;
    return guard(ctx, [&] {
        if (ctx->systems.empty()) throw StateError("no system defined");
        std::vector<const mrfg::CoupledSystem *> ptrs;
        for (auto &s : ctx->systems) {
            try {
                s->check();
            } catch (const std::runtime_error &e) {
                throw ArgError(e.what());
            }
            ptrs.push_back(s.get());
        }
        ctx->dna.reset();
        ctx->batch->setSystems(ptrs);
        ctx->committed = true;
    });
}

int mrfg_update_region(mrfg_ctx *ctx, int sys, int region, const double *diag, const double *lower,
                       const double *upper, const double *source)
{
    if (!ctx) return MRFG_ERR_ARG;
    return guard(ctx, [&] {
        mrfg::Batch &B = committedBatch(ctx);
        mrfg::CoupledSystem &S = systemOf(ctx, sys);
        if (region < 0 || region >= (int)S.regions.size()) throw ArgError("region index out of range");
        B.updateRegion(sys, region, diag, lower, upper, source);
        // the definition follows, so a later commit keeps these values
        mrfg::LduRegion &R = S.regions[region];
        R.diag.assign(diag, diag + R.nCells);
        R.source.assign(source, source + R.nCells);
        if (R.nFaces > 0) {
            R.upper.assign(upper, upper + R.nFaces);
            if (lower) R.lower.assign(lower, lower + R.nFaces);
            else R.lower = R.upper;
        }
    });
}

int mrfg_update_interface(mrfg_ctx *ctx, int sys, int iface, const double *coeff_a, const double *coeff_b)
{
    if (!ctx) return MRFG_ERR_ARG;
    return guard(ctx, [&] {
        mrfg::Batch &B = committedBatch(ctx);
        mrfg::CoupledSystem &S = systemOf(ctx, sys);
        if (iface < 0 || iface >= (int)S.interfaces.size()) throw ArgError("interface index out of range");
        B.updateInterface(sys, iface, coeff_a, coeff_b);
        mrfg::CoupledInterface &I = S.interfaces[iface];
        if (coeff_a) I.a.coeff.assign(coeff_a, coeff_a + I.a.nFaces());
        if (coeff_b) I.b.coeff.assign(coeff_b, coeff_b + I.b.nFaces());
    });
}

int mrfg_dna_solve(mrfg_ctx *ctx, const mrfg_controls *controls, mrfg_dna_perf *perf)
{
    if (!ctx || !controls || !perf) return MRFG_ERR_ARG;
    return guard(ctx, [&] {
        mrfg::Batch &B = committedBatch(ctx);
        std::vector<const mrfg::CoupledSystem *> cases;
        for (auto &s : ctx->systems) {
            if (!s->part) throw ArgError("system " + std::to_string(cases.size()) +
                                         " is not a partitioned case (read one with regionCoupledTemperatureJump / "
                                         "regionCoupledHeatFlux patches)");
            cases.push_back(s.get());
        }
        if (!ctx->dna) ctx->dna.reset(new mrfg::DnaRun(ctx->device, cases, ctx->preferLevels));
        std::vector<mrfg::DnaPerf> out;
        ctx->dna->run(fromControls(controls), out);
        for (size_t q = 0; q < out.size(); ++q) {
            mrfg_dna_perf &p = perf[q];
            p.iterations = out[q].iterations;
            p.converged = out[q].converged;
            p.region_solves = out[q].regionSolves;
            p.solver_iterations = out[q].solverIterations;
            p.jump_residual = out[q].jumpResidual;
            p.flux_residual = out[q].fluxResidual;
            p.seconds = out[q].seconds;
            // the coupled fields become the systems' current fields
            const mrfg::CoupledSystem &S = *ctx->systems[q];
            for (size_t r = 0; r < S.regions.size(); ++r) {
                std::vector<double> psi(S.regions[r].nCells);
                ctx->dna->getPsi((int)q, (int)r, psi.data());
                B.setPsi((int)q, (int)r, psi.data());
            }
        }
    });
}

int mrfg_dna_time_steps(mrfg_ctx *ctx, const mrfg_controls *controls, const char *ddt_scheme, int n_steps,
                        mrfg_dna_perf *perf)
{
    if (!ctx || !controls || !ddt_scheme || !perf || n_steps < 1) return MRFG_ERR_ARG;
    return guard(ctx, [&] {
        const std::string sch = ddt_scheme;
        if (sch != "Euler" && sch != "backward") throw ArgError("ddt scheme must be Euler or backward");
        mrfg::Batch &B = committedBatch(ctx);
        std::vector<const mrfg::CoupledSystem *> cases;
        for (auto &s : ctx->systems) {
            if (!s->part) throw ArgError("system " + std::to_string(cases.size()) + " is not a partitioned case");
            cases.push_back(s.get());
        }
        if (!ctx->dna) ctx->dna.reset(new mrfg::DnaRun(ctx->device, cases, ctx->preferLevels));
        std::vector<mrfg::DnaPerf> out;
        ctx->dna->timeSteps(fromControls(controls), sch == "backward", n_steps, out);
        for (size_t q = 0; q < out.size(); ++q) {
            mrfg_dna_perf &p = perf[q];
            p.iterations = out[q].iterations;
            p.converged = out[q].converged;
            p.region_solves = out[q].regionSolves;
            p.solver_iterations = out[q].solverIterations;
            p.jump_residual = out[q].jumpResidual;
            p.flux_residual = out[q].fluxResidual;
            p.seconds = out[q].seconds;
        }
        for (size_t q = 0; q < ctx->systems.size(); ++q) {
            const mrfg::CoupledSystem &S = *ctx->systems[q];
            for (size_t r = 0; r < S.regions.size(); ++r) {
                std::vector<double> psi(S.regions[r].nCells);
                ctx->dna->getPsi((int)q, (int)r, psi.data());
                B.setPsi((int)q, (int)r, psi.data());
            }
        }
    });
}

int mrfg_dna_history(mrfg_ctx *ctx, int sys, int max_iter, double *jump, double *flux, double *omega, int *n)
{
    if (!ctx || !n || max_iter < 0) return MRFG_ERR_ARG;
    return guard(ctx, [&] {
        if (!ctx->dna) throw StateError("no partitioned solve has run");
        if (sys < 0 || sys >= ctx->dna->nCases()) throw ArgError("system index out of range");
        const auto &h = ctx->dna->history(sys);
        *n = (int)h.size();
        for (int i = 0; i < std::min<int>(max_iter, (int)h.size()); ++i) {
            if (jump) jump[i] = h[i][0];
            if (flux) flux[i] = h[i][1];
            if (omega) omega[i] = h[i][2];
        }
    });
}

int mrfg_solve(mrfg_ctx *ctx, const mrfg_controls *controls, mrfg_perf *perf)
{
    if (!ctx || !controls || !perf) return MRFG_ERR_ARG;
    return guard(ctx, [&] {
        mrfg::Batch &B = committedBatch(ctx);
        std::vector<mrfg::SolvePerf> p;
        B.solve(fromControls(controls), p);
        fillPerf(p, perf);
    });
}

int mrfg_time_steps(mrfg_ctx *ctx, const mrfg_controls *controls, const char *ddt_scheme, int n_steps,
                    mrfg_perf *perf)
{
    if (!ctx || !controls || !ddt_scheme || !perf || n_steps < 1) return MRFG_ERR_ARG;
    return guard(ctx, [&] {
        const std::string s = ddt_scheme;
        if (s != "Euler" && s != "backward") throw ArgError("ddt scheme must be Euler or backward");
        mrfg::Batch &B = committedBatch(ctx);
        std::vector<mrfg::SolvePerf> p;
        B.timeSteps(fromControls(controls), s == "backward", n_steps, p);
        fillPerf(p, perf);
    });
}

int mrfg_solve_dict(mrfg_ctx *ctx, const char *fv_solution, const char *entry, mrfg_perf *perf)
{
    if (!ctx || !perf) return MRFG_ERR_ARG;
    return guard(ctx, [&] {
        mrfg::SolverControls c = controlsFromText(fv_solution, entry);
        mrfg::Batch &B = committedBatch(ctx);
        std::vector<mrfg::SolvePerf> p;
        B.solve(c, p);
        fillPerf(p, perf);
    });
}

int mrfg_parse_controls(const char *fv_solution, const char *entry, mrfg_controls *out, char *err, int err_len)
{
    if (!out) return MRFG_ERR_ARG;
    mrfg_ctx tmp;
    int rc = guard(&tmp, [&] { toControls(controlsFromText(fv_solution, entry), out); });
    if (rc != MRFG_OK) copyErr(tmp.err, err, err_len);
    return rc;
}

int mrfg_reset_psi(mrfg_ctx *ctx)
{
    if (!ctx) return MRFG_ERR_ARG;
    return guard(ctx, [&] { committedBatch(ctx).resetPsi(); });
}

int mrfg_get_psi(mrfg_ctx *ctx, int sys, int region, double *psi)
{
    if (!ctx || !psi) return MRFG_ERR_ARG;
    return guard(ctx, [&] {
        systemOf(ctx, sys);
        committedBatch(ctx).getPsi(sys, region, psi);
    });
}

int mrfg_set_psi(mrfg_ctx *ctx, int sys, int region, const double *psi)
{
    if (!ctx || !psi) return MRFG_ERR_ARG;
    return guard(ctx, [&] {
        mrfg::CoupledSystem &S = systemOf(ctx, sys);
        if (region < 0 || region >= (int)S.regions.size()) throw ArgError("region index out of range");
        S.regions[region].psi.assign(psi, psi + S.regions[region].nCells);
        if (ctx->committed) ctx->batch->setPsi(sys, region, psi);
    });
}

// ---- case builder ----------------------------------------------------------

int mrfg_case_build(const char *dict_text, mrfg_case **out, char *err, int err_len)
{
    if (!dict_text || !out) return MRFG_ERR_ARG;
    *out = nullptr;
    mrfg_ctx tmp;
    auto c = std::make_unique<mrfg_case>();
    int rc = guard(&tmp, [&] { c->sys = mrfg::buildCase(dict_text); });
    if (rc != MRFG_OK) {
        copyErr(tmp.err, err, err_len);
        return rc;
    }
    *out = c.release();
    return MRFG_OK;
}

int mrfg_case_read_foam(const char *case_dir, const char *options, mrfg_case **out, char *err, int err_len)
{
    if (!case_dir || !out) return MRFG_ERR_ARG;
    *out = nullptr;
    mrfg_ctx tmp;
    auto c = std::make_unique<mrfg_case>();
    int rc = guard(&tmp, [&] { c->sys = mrfg::readFoamCase(case_dir, options ? options : ""); });
    if (rc != MRFG_OK) {
        copyErr(tmp.err, err, err_len);
        return rc;
    }
    *out = c.release();
    return MRFG_OK;
}

void mrfg_case_free(mrfg_case *c) { delete c; }

int mrfg_case_sizes(const mrfg_case *c, int *n_regions, int *n_interfaces)
{
    if (!c) return MRFG_ERR_ARG;
    if (n_regions) *n_regions = (int)c->sys.regions.size();
    if (n_interfaces) *n_interfaces = (int)c->sys.interfaces.size();
    return MRFG_OK;
}

int mrfg_case_region_sizes(const mrfg_case *c, int region, int *n_cells, int *n_faces, int *symmetric)
{
    if (!c || region < 0 || region >= (int)c->sys.regions.size()) return MRFG_ERR_ARG;
    const auto &R = c->sys.regions[region];
    if (n_cells) *n_cells = R.nCells;
    if (n_faces) *n_faces = R.nFaces;
    if (symmetric) *symmetric = R.symmetric ? 1 : 0;
    return MRFG_OK;
}

int mrfg_case_region(const mrfg_case *c, int region, int32_t *l_addr, int32_t *u_addr, double *diag, double *lower,
                     double *upper, double *source, double *psi0)
{
    if (!c || region < 0 || region >= (int)c->sys.regions.size()) return MRFG_ERR_ARG;
    const auto &R = c->sys.regions[region];
    auto cp = [](auto *dst, const auto &src) {
        if (dst) std::memcpy(dst, src.data(), src.size() * sizeof(src[0]));
    };
    cp(l_addr, R.lAddr);
    cp(u_addr, R.uAddr);
    cp(diag, R.diag);
    cp(lower, R.lower);
    cp(upper, R.upper);
    cp(source, R.source);
    cp(psi0, R.psi);
    return MRFG_OK;
}

int mrfg_case_region_ddt(const mrfg_case *c, int region, int *transient, double *ddt_coeff, double *diag_base,
                         double *source_base)
{
    if (!c || region < 0 || region >= (int)c->sys.regions.size()) return MRFG_ERR_ARG;
    const auto &R = c->sys.regions[region];
    if (transient) *transient = R.transient() ? 1 : 0;
    if (!R.transient()) return MRFG_OK;
    auto cp = [](double *dst, const std::vector<double> &src) {
        if (dst) std::memcpy(dst, src.data(), src.size() * sizeof(double));
    };
    cp(ddt_coeff, R.ddtCoeff);
    cp(diag_base, R.diagBase);
    cp(source_base, R.sourceBase);
    return MRFG_OK;
}

int mrfg_case_interface_sizes(const mrfg_case *c, int iface, int *region_a, int *region_b, int *n_faces_a,
                              int *n_faces_b, int *n_pairs_a, int *n_pairs_b)
{
    if (!c || iface < 0 || iface >= (int)c->sys.interfaces.size()) return MRFG_ERR_ARG;
    const auto &I = c->sys.interfaces[iface];
    if (region_a) *region_a = I.a.region;
    if (region_b) *region_b = I.b.region;
    if (n_faces_a) *n_faces_a = I.a.nFaces();
    if (n_faces_b) *n_faces_b = I.b.nFaces();
    if (n_pairs_a) *n_pairs_a = (int)I.a.w.size();
    if (n_pairs_b) *n_pairs_b = (int)I.b.w.size();
    return MRFG_OK;
}

int mrfg_case_dna_controls(const mrfg_case *c, int *n_pairs, int *max_couple_iter, double *max_jump_res,
                           double *max_flux_res)
{
    if (!c || !n_pairs) return MRFG_ERR_ARG;
    const auto &P = c->sys.part;
    *n_pairs = P ? (int)P->pairs.size() : 0;
    if (max_couple_iter) *max_couple_iter = P ? P->maxCoupleIter : 0;
    if (max_jump_res) *max_jump_res = P ? P->maxJumpRes : -1.0;
    if (max_flux_res) *max_flux_res = P ? P->maxFluxRes : -1.0;
    return MRFG_OK;
}

namespace {
const mrfg::PartitionedSide *dnaSide(const mrfg_case *c, int pair, int side)
{
    if (!c || !c->sys.part || pair < 0 || pair >= (int)c->sys.part->pairs.size() || side < 0 || side > 1)
        return nullptr;
    const auto &P = c->sys.part->pairs[pair];
    return side ? &P.neu : &P.dir;
}
}  // namespace

int mrfg_case_dna_side_sizes(const mrfg_case *c, int pair, int side, int *region, int *n_faces, int *n_weights,
                             int *acc_type, double *relax, double *K)
{
    const mrfg::PartitionedSide *S = dnaSide(c, pair, side);
    if (!S) return MRFG_ERR_ARG;
    if (region) *region = S->region;
    if (n_faces) *n_faces = S->nFaces();
    if (n_weights) *n_weights = (int)S->w.size();
    if (acc_type) *acc_type = S->accType;
    if (relax) *relax = S->relax;
    if (K) *K = c->sys.part->pairs[pair].K;
    return MRFG_OK;
}

int mrfg_case_dna_side(const mrfg_case *c, int pair, int side, int32_t *fc, double *coeff, double *kb, double *dc,
                       double *value0, int32_t *off, int32_t *idx, double *w)
{
    const mrfg::PartitionedSide *S = dnaSide(c, pair, side);
    if (!S) return MRFG_ERR_ARG;
    auto cp = [](auto *dst, const auto &src) {
        if (dst) std::memcpy(dst, src.data(), src.size() * sizeof(src[0]));
    };
    cp(fc, S->fc); cp(coeff, S->coeff); cp(kb, S->kb); cp(dc, S->dc); cp(value0, S->value0);
    cp(off, S->off); cp(idx, S->idx); cp(w, S->w);
    return MRFG_OK;
}

int mrfg_case_region_source_base(const mrfg_case *c, int region, double *source_base)
{
    if (!c || !source_base || !c->sys.part || region < 0 || region >= (int)c->sys.part->sourceBase.size())
        return MRFG_ERR_ARG;
    const auto &v = c->sys.part->sourceBase[region];
    std::memcpy(source_base, v.data(), v.size() * sizeof(double));
    return MRFG_OK;
}

int mrfg_case_interface(const mrfg_case *c, int iface, int32_t *fc_a, double *coeff_a, int32_t *off_a,
                        int32_t *idx_a, double *w_a, int32_t *fc_b, double *coeff_b, int32_t *off_b, int32_t *idx_b,
                        double *w_b)
{
    if (!c || iface < 0 || iface >= (int)c->sys.interfaces.size()) return MRFG_ERR_ARG;
    const auto &I = c->sys.interfaces[iface];
    auto cp = [](auto *dst, const auto &src) {
        if (dst) std::memcpy(dst, src.data(), src.size() * sizeof(src[0]));
    };
    cp(fc_a, I.a.fc); cp(coeff_a, I.a.coeff); cp(off_a, I.a.off); cp(idx_a, I.a.idx); cp(w_a, I.a.w);
    cp(fc_b, I.b.fc); cp(coeff_b, I.b.coeff); cp(off_b, I.b.off); cp(idx_b, I.b.idx); cp(w_b, I.b.w);
    return MRFG_OK;
}

int mrfg_case_interface_wall(const mrfg_case *c, int iface, double *k_own_a, double *k_nei_a, double *k_own_b,
                             double *k_nei_b)
{
    if (!c || iface < 0 || iface >= (int)c->sys.interfaces.size()) return MRFG_ERR_ARG;
    const auto &I = c->sys.interfaces[iface];
    if (I.a.kOwn.empty() && I.a.nFaces() > 0) return MRFG_ERR_STATE;
    auto cp = [](double *dst, const std::vector<double> &src) {
        if (dst && !src.empty()) std::memcpy(dst, src.data(), src.size() * sizeof(double));
    };
    cp(k_own_a, I.a.kOwn); cp(k_nei_a, I.a.kNei); cp(k_own_b, I.b.kOwn); cp(k_nei_b, I.b.kNei);
    return MRFG_OK;
}

int mrfg_add_case(mrfg_ctx *ctx, const mrfg_case *c, int *sys_id)
{
    if (!ctx || !c) return MRFG_ERR_ARG;
    return guard(ctx, [&] {
        ctx->systems.push_back(std::make_unique<mrfg::CoupledSystem>(c->sys));
        ctx->committed = false;
        if (sys_id) *sys_id = (int)ctx->systems.size() - 1;
    });
}

// ---- kernel-level -----------------------------------------------------------

int mrfg_spmv(mrfg_ctx *ctx, int sys, const double *x, double *y)
{
    if (!ctx || !x || !y) return MRFG_ERR_ARG;
    return guard(ctx, [&] {
        systemOf(ctx, sys);
        committedBatch(ctx).spmv(x, y, sys);
    });
}

int mrfg_dilu_factor(mrfg_ctx *ctx, int sys, double *rD)
{
    if (!ctx || !rD) return MRFG_ERR_ARG;
    return guard(ctx, [&] {
        systemOf(ctx, sys);
        committedBatch(ctx).diluFactor(sys, rD);
    });
}

int mrfg_precondition(mrfg_ctx *ctx, int sys, const double *r, double *w)
{
    if (!ctx || !r || !w) return MRFG_ERR_ARG;
    return guard(ctx, [&] {
        systemOf(ctx, sys);
        committedBatch(ctx).precondition(r, w, sys);
    });
}

int mrfg_tuning(mrfg_ctx *ctx, int *tuned, int *block_threads, int *ell_chunk, double *spmv_ms,
                double *spmv_ms_default)
{
    if (!ctx) return MRFG_ERR_ARG;
    return guard(ctx, [&] {
        const mrfg::Tuning &t = committedBatch(ctx).tuning();
        if (tuned) *tuned = t.tuned ? 1 : 0;
        if (block_threads) *block_threads = 64 * t.slices;
        if (ell_chunk) *ell_chunk = t.ellChunk;
        if (spmv_ms) *spmv_ms = t.bestMs;
        if (spmv_ms_default) *spmv_ms_default = t.defaultMs;
    });
}

int mrfg_set_tuning(mrfg_ctx *ctx, int block_threads, int ell_chunk)
{
    if (!ctx) return MRFG_ERR_ARG;
    return guard(ctx, [&] {
        if (block_threads != 256 && block_threads != 512 && block_threads != 1024)
            throw ArgError("block_threads must be 256, 512 or 1024");
        if (ell_chunk != 2 && ell_chunk != 4 && ell_chunk != 8 && ell_chunk != 16)
            throw ArgError("ell_chunk must be 2, 4, 8 or 16");
        committedBatch(ctx).forceTuning(block_threads / 64, ell_chunk);
    });
}

int mrfg_interface_flux(mrfg_ctx *ctx, int sys, int iface, double *q_a, double *q_b)
{
    if (!ctx || !q_a || !q_b) return MRFG_ERR_ARG;
    return guard(ctx, [&] {
        systemOf(ctx, sys);
        mrfg::Batch &B = committedBatch(ctx);
        if (iface < 0 || iface >= (int)ctx->systems[sys]->interfaces.size())
            throw std::invalid_argument("interface index out of range");
        B.interfaceFlux(sys, iface, q_a, q_b);
    });
}

int mrfg_set_interface_wall(mrfg_ctx *ctx, int sys, int iface, const double *k_own_a, const double *k_nei_a,
                            const double *k_own_b, const double *k_nei_b)
{
    if (!ctx) return MRFG_ERR_ARG;
    return guard(ctx, [&] {
        mrfg::CoupledSystem &S = systemOf(ctx, sys);
        if (iface < 0 || iface >= (int)S.interfaces.size()) throw ArgError("interface index out of range");
        mrfg::CoupledInterface &I = S.interfaces[iface];
        if (I.a.region < 0) throw StateError("set the interface before its wall conductances");
        const bool na = I.a.nFaces() > 0, nb = I.b.nFaces() > 0;
        if ((na && (!k_own_a || !k_nei_a)) || (nb && (!k_own_b || !k_nei_b)))
            throw ArgError("null wall conductance array");
        for (auto side : {std::make_tuple(&I.a, k_own_a, k_nei_a), std::make_tuple(&I.b, k_own_b, k_nei_b)}) {
            mrfg::InterfaceSide &IS = *std::get<0>(side);
            const int n = IS.nFaces();
            for (int f = 0; f < n; ++f)
                if (!(std::get<1>(side)[f] > 0.0) || !(std::get<2>(side)[f] > 0.0))
                    throw ArgError("wall conductances must be positive");
            IS.kOwn.assign(std::get<1>(side), std::get<1>(side) + n);
            IS.kNei.assign(std::get<2>(side), std::get<2>(side) + n);
        }
        if (ctx->committed && ctx->batch) {
            ctx->batch->setInterfaceWall(sys, iface, 0, I.a.kOwn.data(), I.a.kNei.data());
            ctx->batch->setInterfaceWall(sys, iface, 1, I.b.kOwn.data(), I.b.kNei.data());
        }
    });
}

int mrfg_interface_patch_values(mrfg_ctx *ctx, int sys, int iface, int side, double *t_neighbour, double *t_wall)
{
    if (!ctx || !t_neighbour || !t_wall) return MRFG_ERR_ARG;
    return guard(ctx, [&] {
        systemOf(ctx, sys);
        mrfg::Batch &B = committedBatch(ctx);
        if (iface < 0 || iface >= (int)ctx->systems[sys]->interfaces.size())
            throw ArgError("interface index out of range");
        if (side < 0 || side > 1) throw ArgError("side must be 0 (a) or 1 (b)");
        B.interfacePatchValues(sys, iface, side, t_neighbour, t_wall);
    });
}

int mrfg_interface_flux_noalloc(mrfg_ctx *ctx, int sys, int iface, double *q_a, double *q_b)
{
    if (!ctx || !q_a || !q_b) return MRFG_ERR_ARG;
    return guard(ctx, [&] {
        systemOf(ctx, sys);
        mrfg::Batch &B = committedBatch(ctx);
        if (iface < 0 || iface >= (int)ctx->systems[sys]->interfaces.size())
            throw ArgError("interface index out of range");
        if (hipSetDevice(ctx->device) != hipSuccess) throw std::runtime_error("HIP error: hipSetDevice");
        double *scratch = mrfg::fluxScratch(ctx->device);   // taken at mrfg_create; cached per device
        B.interfaceFluxScratch(sys, iface, q_a, q_b, scratch, mrfg::kFluxScratchDoubles);
    });
}

// ---- finite-volume assembly ---------------------------------------------------

int mrfg_assembly_bytes(mrfg_ctx *ctx, double *assemble_bytes, double *place_bytes)
{
    if (!ctx || !assemble_bytes || !place_bytes) return MRFG_ERR_ARG;
    return guard(ctx, [&] { committedBatch(ctx).assemblyBytes(*assemble_bytes, *place_bytes); });
}

int mrfg_set_field(mrfg_ctx *ctx, int sys, int region, const char *name, const double *values, long long n)
{
    if (!ctx || !name || !values || n < 1) return MRFG_ERR_ARG;
    return guard(ctx, [&] {
        systemOf(ctx, sys);
        mrfg::Batch &B = committedBatch(ctx);
        if (!B.hasModel(sys)) throw ArgError("system has no finite-volume model (read a case with mrfg_case_read_foam)");
        B.setField(sys, region, name, values, (size_t)n);
    });
}

int mrfg_assemble(mrfg_ctx *ctx)
{
    if (!ctx) return MRFG_ERR_ARG;
    return guard(ctx, [&] {
        mrfg::Batch &B = committedBatch(ctx);
        bool any = false;
        for (int q = 0; q < B.nSystems(); ++q) any = any || B.hasModel(q);
        if (!any) throw ArgError("no committed system has a finite-volume model (read a case with mrfg_case_read_foam)");
        B.assemble();
    });
}

int mrfg_get_region_matrix(mrfg_ctx *ctx, int sys, int region, double *diag, double *lower, double *upper,
                           double *source)
{
    if (!ctx || !diag || !lower || !upper || !source) return MRFG_ERR_ARG;
    return guard(ctx, [&] {
        const mrfg::CoupledSystem &S = systemOf(ctx, sys);
        if (region < 0 || region >= (int)S.regions.size()) throw ArgError("region index out of range");
        mrfg::Batch &B = committedBatch(ctx);
        if (!B.hasModel(sys)) throw ArgError("system has no finite-volume model");
        B.getRegionMatrix(sys, region, diag, lower, upper, source);
    });
}

int mrfg_get_interface_coeff(mrfg_ctx *ctx, int sys, int iface, int side, double *coeff)
{
    if (!ctx || !coeff || side < 0 || side > 1) return MRFG_ERR_ARG;
    return guard(ctx, [&] {
        const mrfg::CoupledSystem &S = systemOf(ctx, sys);
        if (iface < 0 || iface >= (int)S.interfaces.size()) throw ArgError("interface index out of range");
        mrfg::Batch &B = committedBatch(ctx);
        if (!B.hasModel(sys)) throw ArgError("system has no finite-volume model");
        B.getInterfaceCoeff(sys, iface, side, coeff);
    });
}

int mrfg_precondition_fp32(mrfg_ctx *ctx, int sys, const float *r, float *w)
{
    if (!ctx || !r || !w) return MRFG_ERR_ARG;
    return guard(ctx, [&] {
        systemOf(ctx, sys);
        committedBatch(ctx).precondition32(r, w, sys);
    });
}

int mrfg_precondition_inner(mrfg_ctx *ctx, int sys, int inner_preconditioner, const float *r, float *w)
{
    if (!ctx || !r || !w) return MRFG_ERR_ARG;
    if (inner_preconditioner != MRFG_INNER_TILE_DILU && inner_preconditioner != MRFG_INNER_DILU) return MRFG_ERR_ARG;
    return guard(ctx, [&] {
        systemOf(ctx, sys);
        committedBatch(ctx).precondition32(r, w, sys, inner_preconditioner == MRFG_INNER_TILE_DILU);
    });
}

int mrfg_dilu_tiles(mrfg_ctx *ctx, int sys, int *tile_of_cell)
{
    if (!ctx || !tile_of_cell) return MRFG_ERR_ARG;
    return guard(ctx, [&] {
        systemOf(ctx, sys);
        committedBatch(ctx).tilesOf(sys, tile_of_cell);
    });
}

int mrfg_time_kernel(mrfg_ctx *ctx, int which, int reps, double *ms)
{
    if (!ctx || !ms || reps < 1 || which < 0 || which > 8) return MRFG_ERR_ARG;
    return guard(ctx, [&] { *ms = committedBatch(ctx).timeKernel(which, reps); });
}

int mrfg_trace_dilu(mrfg_ctx *ctx, long long *out, int max_waves, int *n_waves)
{
    if (!ctx || !out || max_waves < 1) return MRFG_ERR_ARG;
    return guard(ctx, [&] {
        int n = committedBatch(ctx).traceDilu(out, max_waves);
        if (n_waves) *n_waves = n;
    });
}

int mrfg_layout_info(mrfg_ctx *ctx, int sys, long long *info)
{
    if (!ctx || !info) return MRFG_ERR_ARG;
    return guard(ctx, [&] {
        systemOf(ctx, sys);
        const mrfg::SystemLayout &L = committedBatch(ctx).layout(sys);
        info[0] = L.nCells;
        info[1] = L.nFaces;
        info[2] = L.nPairs;
        info[3] = L.nPos;
        info[4] = L.nSlices;
        info[5] = L.nWaves;
    });
}

int mrfg_copy_probe(mrfg_ctx *ctx, long long bytes, int reps, double *ms)
{
    if (!ctx || !ms || bytes < 16 || reps < 1) return MRFG_ERR_ARG;
    return guard(ctx, [&] {
        hipError_t e = mrfg::copyProbe(ctx->device, (size_t)bytes, reps, ms);
        if (e != hipSuccess) throw std::runtime_error(std::string("HIP error: ") + hipGetErrorString(e));
    });
}

int mrfg_bytes(mrfg_ctx *ctx, int sys, int value_size, double *b_iter, double *b_spmv, double *b_dilu)
{
    if (!ctx || (value_size != 4 && value_size != 8)) return MRFG_ERR_ARG;
    return guard(ctx, [&] {
        systemOf(ctx, sys);
        mrfg::Batch &B = committedBatch(ctx);
        if (b_iter) *b_iter = B.bytesIter(sys, value_size);
        if (b_spmv) *b_spmv = B.bytesSpmv(sys, value_size);
        if (b_dilu) *b_dilu = B.bytesDilu(sys, value_size);
    });
}

/* ---- serving ----------------------------------------------------------- */

struct mrfg_server {
    std::unique_ptr<mrfg::Server> srv;
    std::string err;
};

int mrfg_server_create(mrfg_server **out, int device, long long hbm_budget_bytes, int n_workers, int max_batch,
                       int n_prep, const char *fv_solution, const char *entry, int keep_fields)
{
    if (!out || !fv_solution || !entry || hbm_budget_bytes <= 0 || n_workers < 1 || max_batch < 1 || n_prep < 0)
        return MRFG_ERR_ARG;
    *out = nullptr;
    try {
        mrfg::Dict d = mrfg::parseDict(fv_solution);
        mrfg::SolverControls c = mrfg::parseControls(d.subDict("solvers").subDict(entry));
        auto *s = new mrfg_server;
        s->srv.reset(new mrfg::Server(device, (size_t)hbm_budget_bytes, n_workers, max_batch, c, keep_fields != 0,
                                      n_prep));
        *out = s;
        return MRFG_OK;
    } catch (const mrfg::DictError &) {
        return MRFG_ERR_DICT;
    } catch (const std::exception &) {
        return MRFG_ERR_DEVICE;
    }
}

void mrfg_server_destroy(mrfg_server *s) { delete s; }

int mrfg_server_submit(mrfg_server *s, const char *case_dict, int *request_id)
{
    if (!s || !case_dict || !request_id) return MRFG_ERR_ARG;
    *request_id = s->srv->submit(case_dict);
    return MRFG_OK;
}

int mrfg_server_wait_all(mrfg_server *s)
{
    if (!s) return MRFG_ERR_ARG;
    s->srv->waitAll();
    return MRFG_OK;
}

int mrfg_server_result(mrfg_server *s, int id, mrfg_perf *perf, double *times, int *batch_size)
{
    if (!s) return MRFG_ERR_ARG;
    mrfg::ServeResult R;   // a copy taken under the server's lock
    if (!s->srv->result(id, R)) return MRFG_ERR_ARG;
    if (R.status == -1) {
        s->err = R.error;
        return MRFG_ERR_DEVICE;
    }
    if (R.status != 2) return MRFG_ERR_STATE;
    if (perf) fillPerf(std::vector<mrfg::SolvePerf>{R.perf}, perf);
    if (times) {
        times[0] = R.tSubmit;
        times[1] = R.tAdmit;
        times[2] = R.tDone;
    }
    if (batch_size) *batch_size = R.batchSize;
    return MRFG_OK;
}

int mrfg_server_times(mrfg_server *s, int id, double *times)
{
    if (!s || !times) return MRFG_ERR_ARG;
    mrfg::ServeResult R;
    if (!s->srv->result(id, R)) return MRFG_ERR_ARG;
    if (R.status != 2) return MRFG_ERR_STATE;
    times[0] = R.tSubmit;
    times[1] = R.tReady;
    times[2] = R.tAdmit;
    times[3] = R.tSolve;
    times[4] = R.tDone;
    return MRFG_OK;
}

int mrfg_server_get_psi(mrfg_server *s, int id, int region, double *psi)
{
    if (!s || !psi) return MRFG_ERR_ARG;
    return s->srv->copyPsi(id, region, psi) ? MRFG_OK : MRFG_ERR_STATE;
}

const char *mrfg_server_last_error(const mrfg_server *s) { return s ? s->err.c_str() : ""; }

}  // extern "C"
