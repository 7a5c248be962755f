// Batched 2D fitting: many independent Nelder-Mead fits advanced in lockstep.
//
// The simplex is the one of SasCalc_SimplexFit2D::doSimplexFit2D
// (Source/sc_simplexfit2d.cpp:254-611) with its trial step amotry (:221-251) and the bounds
// rule of fehlerquadratsumme (:25-44), written as a resumable state machine: instead of
// calling the objective, a fit posts the points it needs and resumes when their values
// arrive.  The driver collects the points of every fit, prepares each parameter set on host
// threads, stages them all with csx_fit_batch_stage and evaluates them in one
// csx_fit_batch_run launch.  Independent evaluations of one fit (the initial vertices and a
// shrink) are posted together; the order of those evaluations does not change any value.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <memory>
#include <cmath>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "csx.h"
#include "csx_prep.h"

namespace {

constexpr double kOutOfBounds = 1e20;

class NelderMead {
public:
    // m: length of the parameter vector (the reference's mmax: every entry of the fit's
    // parameter hash); nf: the first nf entries are fitted (numFit), the rest are fixed.
    void init(int m, int nf, int mode, const double* x0, const double* lo, const double* hi, double stp, int maxit,
              double ftol) {
        m_ = m;
        nf_ = nf;
        ref_ = mode == CSX_NM_REFERENCE;
        maxit_ = maxit;
        ftol_ = ftol;
        std::memcpy(lo_, lo, m * sizeof(double));
        std::memcpy(hi_, hi, m * sizeof(double));
        std::memcpy(x0_, x0, m * sizeof(double));
        std::memset(p_, 0, sizeof p_);
        std::memset(y_, 0, sizeof y_);
        // The reference sizes the steps and the trial factor by mmax (:343-346, :231); the
        // corrected simplex by the number of fitted parameters.
        d_ = ref_ ? m : nf;
        const double rootn1 = std::sqrt(d_ + 1.0) - 1.0, mroot2 = 1.0 / (d_ * std::sqrt(2.0));
        double pp[CSX_NM_MAXPAR], qv[CSX_NM_MAXPAR];
        for (int i = 0; i < m; ++i) {
            double step = 0.0;
            if (i < nf) {
                if (x0[i] == 0) step = (lo[i] == 0) ? stp : lo[i] * stp;
                else step = x0[i] * stp;
            }
            pp[i] = step * (rootn1 + d_) * mroot2;
            qv[i] = step * rootn1 * mroot2;
            p_[0][i] = x0[i];
        }
        // Reference: the clearing loop of :384-389 runs its inner loop over i, so it zeroes
        // the first entry of rows 0..29, vertex 0's included, after the vertex was set.
        // y[0] is still evaluated at the start values (:392).
        if (ref_)
            for (int i = 0; i < CSX_NM_MAXPAR; ++i) p_[i][0] = 0.0;
        for (int i = 1; i <= d_; ++i) {
            for (int j = 0; j < m; ++j) p_[i][j] = p_[0][j] + qv[j];
            p_[i][i - 1] = p_[i][i - 1] + pp[i - 1];
        }
        phase_ = kInit;
        rep_ = 0;
        evals_ = 0;
        done_ = false;
        // y[0] at x0, then vertices 1..d; the reference evaluates up to vertex nf+1 and copies
        // y[nf+1] into the rest (:411-424).  The values do not depend on each other, so they
        // are posted together.
        std::vector<const double*> pts{x0_};
        nev0_ = std::min(d_, nf_ + 1);
        for (int i = 1; i <= nev0_; ++i) pts.push_back(p_[i]);
        post(pts);
        drain();
    }

    bool done() const { return done_; }
    int npending() const { return npend_; }
    const double* pending(int k) const { return pts_[k]; }

    // Values of the pending points (in posting order); advances to the next request.
    // nexact: how many of them the exact objective evaluated (the others were screened out).
    void deliver(const double* f, int nexact = -1) {
        evals_ += nexact < 0 ? npend_ : nexact;
        double v[CSX_NM_MAXPAR + 2];
        for (int k = 0; k < nall_; ++k) {
            v[k] = kOutOfBounds;
            spv_[k] = kNoScore;
        }
        for (int w = 0; w < npend_; ++w) {
            v[slot_[w]] = f[w];
            spv_[slot_[w]] = sp_[w];
        }
        step(v);
        drain();
    }

    // Surrogate screening (no reference counterpart).  Scores of the pending points (lower is
    // better), in posting order, before their exact evaluation.
    void set_scores(const double* sc) {
        for (int w = 0; w < npend_; ++w) sp_[w] = sc[w];
    }
    // Pending point w is a trial point whose score exceeds the worst vertex score of the
    // simplex by more than the factor (1 + margin): the simplex would most likely not accept
    // it, so it takes the value of a rejected point (1e20) without an exact evaluation.
    // Initial and shrink vertices are always evaluated: the simplex keeps them.
    bool prunable(int w, double margin) const {
        if (phase_ != kReflect && phase_ != kExpand && phase_ != kContract) return false;
        double worst = -1.0;
        for (int i = 0; i <= nf_; ++i) {
            if (!(sv_[i] < kNoScore)) return false;  // a vertex without a score: no basis
            worst = std::max(worst, sv_[i]);
        }
        return sp_[w] > worst * (1.0 + margin);
    }

    void result(csx_nm_result& r) const {
        for (int j = 0; j < m_; ++j) r.x[j] = p_[ilo_][j];
        r.f = y_[ilo_];
        r.iters = rep_;
        r.evals = evals_;
    }

private:
    enum Phase { kInit, kReflect, kExpand, kContract, kShrink };

    // Post points.  Points outside the bounds are not evaluated: they get 1e20, as the
    // reference's objective returns (:25-44).
    void post(const std::vector<const double*>& xs) {
        nall_ = (int)xs.size();
        npend_ = 0;
        for (int k = 0; k < nall_; ++k) {
            if (out_of_bounds(xs[k])) continue;
            std::memcpy(pts_[npend_], xs[k], m_ * sizeof(double));
            slot_[npend_++] = k;
        }
        resolve_now_ = npend_ == 0;
    }

    // A post with no point in bounds is answered at once.
    void drain() {
        while (resolve_now_ && !done_) {
            resolve_now_ = false;
            double v[CSX_NM_MAXPAR + 2];
            for (int k = 0; k < nall_; ++k) {
                v[k] = kOutOfBounds;
                spv_[k] = kNoScore;
            }
            npend_ = 0;
            step(v);
        }
    }

    // Only the fitted entries are checked (enabled[j], :27-42); bounds 0..360 are exempt.
    bool out_of_bounds(const double* x) const {
        for (int j = 0; j < nf_; ++j) {
            if (lo_[j] == 0.0 && hi_[j] == 360.0) continue;
            if (x[j] < lo_[j] || x[j] > hi_[j]) return true;
        }
        return false;
    }

    // amotry (Source/sc_simplexfit2d.cpp:221-251): trial point of factor fac.  The reference
    // copies every entry of psum first, so its trial points carry the fixed entries at their
    // column sums (:226); the corrected simplex keeps them at their start values.
    void trial(double fac) {
        const double f1 = (1.0 - fac) / d_, f2 = f1 - fac;
        for (int j = 0; j < m_; ++j) pt_[j] = ref_ ? psum_[j] : x0_[j];
        for (int j = 0; j < nf_; ++j) pt_[j] = psum_[j] * f1 - p_[ihi_][j] * f2;
    }
    void accept_if_better(double yt) {
        if (yt < y_[ihi_]) {
            y_[ihi_] = yt;
            sv_[ihi_] = spv_[0];
            for (int j = 0; j < nf_; ++j) {
                psum_[j] += pt_[j] - p_[ihi_][j];
                p_[ihi_][j] = pt_[j];
            }
        }
    }
    void sum_rows() {
        for (int j = 0; j < m_; ++j) {
            double s = 0.0;
            for (int i = 0; i <= nf_; ++i) s += p_[i][j];
            psum_[j] = s;
        }
    }

    // Start of one simplex iteration (the body of the reference's `rep` loop, :440-600).
    void iterate() {
        if (rep_ >= maxit_) { finish(); return; }
        const int mpts = nf_ + 1;
        ilo_ = 0;
        if (y_[0] > y_[1]) { ihi_ = 0; inhi_ = 1; }
        else { ihi_ = 1; inhi_ = 0; }
        for (int i = 0; i < mpts; ++i) {
            if (y_[i] < y_[ilo_]) ilo_ = i;
            if (y_[i] > y_[ihi_]) { inhi_ = ihi_; ihi_ = i; }
            else if (y_[i] > y_[inhi_]) { if (i != ihi_) inhi_ = i; }
        }
        if (std::isnan(y_[ihi_]) || std::isnan(y_[ilo_]) || std::fabs(y_[ihi_]) + std::fabs(y_[ilo_]) == 0) {
            finish();
            return;
        }
        const double rtol = 2.0 * std::fabs(y_[ihi_] - y_[ilo_]) / (std::fabs(y_[ihi_]) + std::fabs(y_[ilo_]));
        if (rtol <= ftol_) { finish(); return; }
        trial(-1.0);  // reflection, alpha = 1
        phase_ = kReflect;
        post({pt_});
    }

    void next_iteration() {
        ++rep_;
        iterate();
    }

    void step(const double* v) {
        switch (phase_) {
        case kInit:
            y_[0] = v[0];
            sv_[0] = spv_[0];
            for (int i = 1; i <= nev0_; ++i) {
                y_[i] = v[i];
                sv_[i] = spv_[i];
            }
            for (int i = nev0_ + 1; i <= d_; ++i) {
                y_[i] = y_[nf_ + 1];
                sv_[i] = sv_[nf_ + 1];
            }
            sum_rows();
            rep_ = 0;
            iterate();
            return;
        case kReflect: {
            const double yt = v[0];
            accept_if_better(yt);
            if (yt <= y_[ilo_]) {
                trial(2.0);  // expansion, gamma = 2
                phase_ = kExpand;
                post({pt_});
            } else if (yt >= y_[inhi_]) {
                ysave_ = y_[ihi_];
                trial(0.5);  // contraction, beta = 0.5
                phase_ = kContract;
                post({pt_});
            } else {
                next_iteration();
            }
            return;
        }
        case kExpand:
            accept_if_better(v[0]);
            next_iteration();
            return;
        case kContract: {
            const double yt = v[0];
            accept_if_better(yt);
            if (yt >= ysave_) {
                // Shrink towards the best vertex over every entry (:566-582).
                std::vector<const double*> xs;
                for (int i = 0; i <= nf_; ++i) {
                    if (i == ilo_) continue;
                    for (int j = 0; j < m_; ++j) p_[i][j] = 0.5 * (p_[i][j] + p_[ilo_][j]);
                    xs.push_back(p_[i]);
                }
                phase_ = kShrink;
                post(xs);
            } else {
                next_iteration();
            }
            return;
        }
        case kShrink: {
            int k = 0;
            for (int i = 0; i <= nf_; ++i)
                if (i != ilo_) {
                    sv_[i] = spv_[k];
                    y_[i] = v[k++];
                }
            sum_rows();
            next_iteration();
            return;
        }
        }
    }

    void finish() {
        done_ = true;
        npend_ = 0;
        nall_ = 0;
    }

    int m_ = 0, nf_ = 0, d_ = 0, nev0_ = 0, maxit_ = 0;
    bool ref_ = true;
    double ftol_ = 0.0, lo_[CSX_NM_MAXPAR], hi_[CSX_NM_MAXPAR], x0_[CSX_NM_MAXPAR];
    double p_[CSX_NM_MAXPAR + 1][CSX_NM_MAXPAR], y_[CSX_NM_MAXPAR + 1], psum_[CSX_NM_MAXPAR], pt_[CSX_NM_MAXPAR];
    double ysave_ = 0.0;
    int ilo_ = 0, ihi_ = 0, inhi_ = 0, rep_ = 0, evals_ = 0;
    Phase phase_ = kInit;
    bool done_ = false, resolve_now_ = false;
    int nall_ = 0, npend_ = 0, slot_[CSX_NM_MAXPAR + 2];
    double pts_[CSX_NM_MAXPAR + 2][CSX_NM_MAXPAR];
    // surrogate scores: of the pending points (posting order and by slot) and of the vertices
    static constexpr double kNoScore = HUGE_VAL;
    double sp_[CSX_NM_MAXPAR + 2] = {}, spv_[CSX_NM_MAXPAR + 2] = {}, sv_[CSX_NM_MAXPAR + 1] = {};
};

// Set one free parameter: a key may name several parameters separated by commas, which
// all take the value (a = b of a tetragonal cell, for example).
void set_tied(csx::ParamSet& ps, const char* key, double v) {
    std::string k;
    for (const char* c = key;; ++c) {
        if (*c == ',' || *c == 0) {
            if (!k.empty()) ps.set(k, v);
            k.clear();
            if (*c == 0) break;
        } else {
            k += *c;
        }
    }
}

// Vector length, fitted count and mode: at least one fitted entry, and the reference's arrays
// hold at most map = 30 entries (mmax < nap = 31, Source/sc_simplexfit2d.h:113-114, :318).
bool nm_args_ok(int npar, int nfree, int mode) {
    return npar >= 1 && npar <= CSX_NM_MAXPAR && nfree >= 1 && nfree <= npar &&
           (mode == CSX_NM_REFERENCE || mode == CSX_NM_CORRECTED);
}

// Run f over [0, n) on up to `threads` host threads.
template <class F>
void parallel_for(int n, int threads, F&& f) {
    threads = std::max(1, std::min(threads, n));
    if (threads == 1) {
        for (int i = 0; i < n; ++i) f(i);
        return;
    }
    std::atomic<int> next{0};
    std::vector<std::thread> pool;
    for (int t = 0; t < threads; ++t)
        pool.emplace_back([&] {
            for (int i = next++; i < n; i = next++) f(i);
        });
    for (auto& th : pool) th.join();
}

}  // namespace

extern "C" {

int csx_nm_run_cb(int npar, int nfree, int mode, csx_objective f, void* user, const double* x0, const double* lo,
                  const double* hi, double stp, int maxit, double ftol, csx_nm_result* out) {
    if (!nm_args_ok(npar, nfree, mode) || !f || !x0 || !lo || !hi || !out) return CSX_E_ARG;
    NelderMead nm;
    nm.init(npar, nfree, mode, x0, lo, hi, stp, maxit, ftol);
    std::vector<double> vals;
    while (!nm.done()) {
        vals.resize(nm.npending());
        for (int k = 0; k < nm.npending(); ++k) vals[k] = f(user, nm.pending(k));
        nm.deliver(vals.data());
    }
    nm.result(*out);
    return CSX_OK;
}

}  // extern "C"

namespace {

// Evaluator of batched points: a point is (fit, parameter vector); one call prepares every
// point on host threads (prepareCalculation + the host part of staging), commits them and
// evaluates them in one batched launch.  Staging slots are kept across calls.
class PointEval {
public:
    PointEval(csx_ctx* ctx, const csx_nm_fit* fits, const std::vector<const csx::ParamSet*>& base, int npar,
              const char* const* keys, int precision, int threads)
        : ctx_(ctx), fits_(fits), base_(base), npar_(npar), keys_(keys), prec_(precision), threads_(threads) {}

    struct Point {
        int fit;
        const double* x;
    };

    int run(const std::vector<Point>& pts, double* sums) {
        if (int rc = launch(pts)) return rc;
        return finish(sums);
    }

    // Prepare and stage the points, then enqueue their evaluation (no wait).
    int launch(const std::vector<Point>& pts) {
        using clk = std::chrono::steady_clock;
        auto ms = [](clk::time_point a, clk::time_point b) { return std::chrono::duration<double, std::milli>(b - a).count(); };
        const int n = (int)pts.size();
        while ((int)slots_.size() < n) slots_.push_back(std::make_unique<Slot>());
        errs_.assign(n, std::string());
        preps_.resize(std::max(preps_.size(), (size_t)n));
        const auto t0 = clk::now();
        // lattice sets: the device builds the reflection tables (csx_set_device_prep)
        const bool dev = csx_ctx_device_prep(ctx_) != 0;
        parallel_for(n, threads_, [&](int r) {
            csx::ParamSet ps = *base_[pts[r].fit];
            for (int j = 0; j < npar_; ++j) set_tied(ps, keys_[j], pts[r].x[j]);
            csx::PreparedSet& P = preps_[r];
            const bool rows = dev && ps.ltype != 12 && ps.peakidx + 1 == 2;  // Gaussian peaks
            errs_[r] = csx::prepare(ps, P, rows);
            if (!errs_[r].empty()) return;
            const csx_nm_fit& F = fits_[pts[r].fit];
            if (rows) {
                csx_stage_build_rows(slots_[r]->p, P, F.w, F.h, &F.mask, F.d_data);
                return;
            }
            csx_stage_build(slots_[r]->p, &P.s, P.carr.data(), P.latpar1.empty() ? nullptr : P.latpar1.data(),
                            P.latpar2.empty() ? nullptr : P.latpar2.data(),
                            P.latpar3.empty() ? nullptr : P.latpar3.data(), F.w, F.h, &F.mask, F.d_data);
        });
        const auto t1 = clk::now();
        // A point whose preparation fails is not evaluated: it scores 1e20, the value
        // doFitCalculation returns on failure (Source/sc_calc_generic_cpu.cpp:1341).
        std::vector<const csx_staged*> built;
        ok_.clear();
        for (int r = 0; r < n; ++r)
            if (errs_[r].empty()) {
                built.push_back(slots_[r]->p);
                ok_.push_back(r);
            }
        n_ = n;
        auto t2 = clk::now();
        if (!built.empty()) {
            int rc = csx_fit_batch_begin(ctx_);
            if (rc) return rc;
            rc = csx_stage_commit_many(ctx_, (int)built.size(), built.data(), threads_);
            if (rc) return rc;
            t2 = clk::now();
            rc = csx_fit_batch_launch(ctx_, prec_);
            if (rc) return rc;
        }
        const auto t3 = clk::now();
        t_prep += ms(t0, t1);
        t_stage += ms(t1, t2);
        t_launch += ms(t2, t3);
        return CSX_OK;
    }

    // Wait for the enqueued points; sums in point order.
    int finish(double* sums) {
        using clk = std::chrono::steady_clock;
        const auto t0 = clk::now();
        for (int r = 0; r < n_; ++r) sums[r] = kOutOfBounds;
        if (ok_.empty()) return CSX_OK;
        cnts_.resize(ok_.size());
        got_.resize(ok_.size());
        const int rc = csx_fit_batch_wait(ctx_, got_.data(), cnts_.data());
        for (size_t k = 0; k < ok_.size(); ++k) sums[ok_[k]] = got_[k];
        t_run += std::chrono::duration<double, std::milli>(clk::now() - t0).count();
        return rc;
    }

    double t_prep = 0, t_stage = 0, t_launch = 0, t_run = 0;

private:
    struct Slot {
        csx_staged* p = csx_staged_new();
        ~Slot() { csx_staged_free(p); }
        Slot() = default;
        Slot(const Slot&) = delete;
        Slot& operator=(const Slot&) = delete;
    };
    csx_ctx* ctx_;
    const csx_nm_fit* fits_;
    const std::vector<const csx::ParamSet*>& base_;
    int npar_;
    const char* const* keys_;
    int prec_, threads_;
    std::vector<std::unique_ptr<Slot>> slots_;
    std::vector<csx::PreparedSet> preps_;
    std::vector<std::string> errs_;
    std::vector<int64_t> cnts_;
    std::vector<int> ok_;
    std::vector<double> got_;
    int n_ = 0;
};

int check_fits(int nfits, const csx_nm_fit* fits, std::vector<const csx::ParamSet*>& base) {
    base.resize(nfits);
    for (int i = 0; i < nfits; ++i) {
        base[i] = csx_engine_paramset(fits[i].base);
        if (!base[i]) return CSX_E_ARG;
    }
    return CSX_OK;
}

}  // namespace

extern "C" {

int csx_nm_batch_screened(csx_ctx* ctx, int nfits, const csx_nm_fit* fits, int npar, int nfree, int mode,
                          const char* const* keys, double stp, int maxit, double ftol, int precision, int threads,
                          const csx_nm_screen_opts* scr, csx_nm_result* out, csx_nm_stats* stats) {
    if (!ctx || nfits < 1 || !fits || !nm_args_ok(npar, nfree, mode) || !keys || !out) return CSX_E_ARG;
    if (scr && scr->fn && !(scr->margin >= 0.0)) return CSX_E_ARG;
    const bool screening = scr && scr->fn;
    std::vector<const csx::ParamSet*> base;
    if (check_fits(nfits, fits, base)) return CSX_E_ARG;
    std::vector<NelderMead> nm(nfits);
    for (int i = 0; i < nfits; ++i) nm[i].init(npar, nfree, mode, fits[i].x0, fits[i].lo, fits[i].hi, stp, maxit, ftol);
    if (threads <= 0) threads = (int)std::max(1u, std::thread::hardware_concurrency());
    // Two groups of fits on two contexts (two streams): while the device evaluates one group's
    // round, the host prepares the other's.  Small batches use one group.
    const int ngroups = nfits >= 32 ? 2 : 1;
    csx_ctx* ctx2 = nullptr;
    if (ngroups == 2 && csx_create_like(ctx, &ctx2) != CSX_OK) return CSX_E_HIP;
    struct Group {
        int f0, f1;
        std::unique_ptr<PointEval> ev;
        std::vector<PointEval::Point> pts;  // every posted point of the round
        std::vector<char> keep;             // evaluated exactly (not screened out)
        std::vector<PointEval::Point> exact;
        std::vector<double> sums;
        bool inflight = false, done = false;
    };
    std::vector<Group> gs(ngroups);
    for (int g = 0; g < ngroups; ++g) {
        gs[g].f0 = (int)((int64_t)nfits * g / ngroups);
        gs[g].f1 = (int)((int64_t)nfits * (g + 1) / ngroups);
        gs[g].ev = std::make_unique<PointEval>(g == 0 ? ctx : ctx2, fits, base, npar, keys, precision, threads);
    }
    int64_t rounds = 0, evals = 0, pruned = 0;
    double screen_ms = 0.0;
    std::vector<int32_t> sfit;
    std::vector<double> sx, sscore;
    int rc = CSX_OK;
    for (int live = ngroups; live > 0 && rc == CSX_OK;) {
        for (auto& G : gs) {
            if (G.done) continue;
            if (G.inflight) {  // deliver the group's previous round
                std::vector<double> got(G.exact.size());
                if ((rc = G.ev->finish(got.data()))) break;
                G.inflight = false;
                G.sums.assign(G.pts.size(), kOutOfBounds);  // screened-out points: rejected
                for (size_t p = 0, e = 0; p < G.pts.size(); ++p)
                    if (G.keep[p]) G.sums[p] = got[e++];
                size_t r = 0;
                for (int i = G.f0; i < G.f1; ++i) {
                    const int np = nm[i].npending();
                    if (np == 0) continue;
                    int nexact = 0;
                    for (int k = 0; k < np; ++k) nexact += G.keep[r + k];
                    nm[i].deliver(G.sums.data() + r, nexact);
                    r += np;
                }
            }
            G.pts.clear();
            for (int i = G.f0; i < G.f1; ++i)
                for (int k = 0; k < nm[i].npending(); ++k) G.pts.push_back({i, nm[i].pending(k)});
            if (G.pts.empty()) {
                G.done = true;
                --live;
                continue;
            }
            G.keep.assign(G.pts.size(), 1);
            if (screening) {
                // surrogate scores of every posted point, then the trial points the simplex
                // would most likely reject are screened out of the exact evaluation
                const auto t0 = std::chrono::steady_clock::now();
                const size_t n = G.pts.size();
                sfit.resize(n);
                sx.resize(n * npar);
                sscore.assign(n, 0.0);
                for (size_t p = 0; p < n; ++p) {
                    sfit[p] = G.pts[p].fit;
                    std::memcpy(&sx[p * npar], G.pts[p].x, npar * sizeof(double));
                }
                if (scr->fn(scr->user, (int)n, sfit.data(), sx.data(), sscore.data()) != 0) {
                    rc = CSX_E_STATE;
                    break;
                }
                size_t r = 0;
                for (int i = G.f0; i < G.f1; ++i) {
                    const int np = nm[i].npending();
                    if (np == 0) continue;
                    nm[i].set_scores(sscore.data() + r);
                    for (int k = 0; k < np; ++k)
                        if (nm[i].prunable(k, scr->margin)) {
                            G.keep[r + k] = 0;
                            ++pruned;
                        }
                    r += np;
                }
                screen_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
            }
            G.exact.clear();
            for (size_t p = 0; p < G.pts.size(); ++p)
                if (G.keep[p]) G.exact.push_back(G.pts[p]);
            ++rounds;
            evals += (int64_t)G.exact.size();
            if ((rc = G.ev->launch(G.exact))) break;
            G.inflight = true;
        }
    }
    double tp = 0, ts = 0, tl = 0, tr = 0;
    for (auto& G : gs) {
        if (G.inflight) {  // an error left a launch in flight: drain it
            std::vector<double> got(G.exact.size());
            (void)G.ev->finish(got.data());
        }
        tp += G.ev->t_prep;
        ts += G.ev->t_stage;
        tl += G.ev->t_launch;
        tr += G.ev->t_run;
    }
    gs.clear();
    if (ctx2) csx_destroy(ctx2);
    if (rc) return rc;
    for (int i = 0; i < nfits; ++i) nm[i].result(out[i]);
    if (stats) {
        stats->rounds = rounds;
        stats->evals = evals;
        stats->prep_ms = tp;
        stats->stage_ms = ts;
        stats->run_ms = tr;
        stats->pruned = pruned;
        stats->screen_ms = screen_ms;
        stats->launch_ms = tl;
    }
    return CSX_OK;
}

int csx_nm_batch(csx_ctx* ctx, int nfits, const csx_nm_fit* fits, int npar, int nfree, int mode,
                 const char* const* keys, double stp, int maxit, double ftol, int precision, int threads,
                 csx_nm_result* out, csx_nm_stats* stats) {
    return csx_nm_batch_screened(ctx, nfits, fits, npar, nfree, mode, keys, stp, maxit, ftol, precision, threads,
                                 nullptr, out, stats);
}

int csx_nm_eval(csx_ctx* ctx, int nfits, const csx_nm_fit* fits, int npar, const char* const* keys, int nv,
                const double* x, int precision, int threads, double* out) {
    if (!ctx || nfits < 1 || !fits || npar < 1 || npar > CSX_NM_MAXPAR || !keys || nv < 1 || !x || !out)
        return CSX_E_ARG;
    std::vector<const csx::ParamSet*> base;
    if (check_fits(nfits, fits, base)) return CSX_E_ARG;
    if (threads <= 0) threads = (int)std::max(1u, std::thread::hardware_concurrency());
    std::vector<PointEval::Point> pts;
    std::vector<size_t> where;
    for (int i = 0; i < nfits; ++i)
        for (int v = 0; v < nv; ++v) {
            const double* xv = x + ((size_t)i * nv + v) * npar;
            bool in = true;
            for (int j = 0; j < npar; ++j)  // the simplex's bounds rule, angles 0..360 exempt
                in = in && ((fits[i].lo[j] == 0.0 && fits[i].hi[j] == 360.0) ||
                            (xv[j] >= fits[i].lo[j] && xv[j] <= fits[i].hi[j]));
            out[(size_t)i * nv + v] = kOutOfBounds;  // the objective's value outside the bounds
            if (in) {
                pts.push_back({i, xv});
                where.push_back((size_t)i * nv + v);
            }
        }
    if (pts.empty()) return CSX_OK;
    // Chunks of points alternate between two contexts, as the rounds of csx_nm_batch do:
    // the host prepares one chunk while the device evaluates the other, and the staging
    // memory stays at the size of a chunk.
    constexpr size_t kChunk = 256;
    const size_t nchunks = (pts.size() + kChunk - 1) / kChunk;
    csx_ctx* ctx2 = nullptr;
    if (nchunks > 1 && csx_create_like(ctx, &ctx2) != CSX_OK) return CSX_E_HIP;
    std::unique_ptr<PointEval> ev[2];
    ev[0] = std::make_unique<PointEval>(ctx, fits, base, npar, keys, precision, threads);
    if (ctx2) ev[1] = std::make_unique<PointEval>(ctx2, fits, base, npar, keys, precision, threads);
    std::vector<double> sums(pts.size());
    std::vector<PointEval::Point> part[2];
    size_t at[2] = {0, 0};
    bool busy[2] = {false, false};
    int rc = CSX_OK;
    for (size_t c = 0; c < nchunks && rc == CSX_OK; ++c) {
        const int e = ctx2 ? (int)(c & 1) : 0;
        if (busy[e]) {
            rc = ev[e]->finish(sums.data() + at[e]);
            busy[e] = false;
            if (rc) break;
        }
        at[e] = c * kChunk;
        part[e].assign(pts.begin() + at[e], pts.begin() + std::min(pts.size(), at[e] + kChunk));
        rc = ev[e]->launch(part[e]);
        busy[e] = rc == CSX_OK;
    }
    for (int e = 0; e < 2; ++e)
        if (busy[e]) {
            const int r2 = ev[e]->finish(sums.data() + at[e]);
            if (rc == CSX_OK) rc = r2;
        }
    ev[0].reset();
    ev[1].reset();
    if (ctx2) csx_destroy(ctx2);
    if (rc) return rc;
    for (size_t k = 0; k < pts.size(); ++k) out[where[k]] = sums[k];
    return CSX_OK;
}

}  // extern "C"
