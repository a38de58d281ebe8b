// GMRES drivers, the potential-table and batch objects, and the C ABI of libmultiscat_amd.
//
// Solve structure (reference: lib/multiscat_fortran/fortran/multiscat_gmres.f90:452-618):
//   Gauss-Seidel split   A = L + U in channel order; GMRES on (I + L^-1 U) psi = L^-1 s,
//                        left-preconditioned by (I - L^-1 U) when neumann is set.
//   Block Jacobi         A = D + V; GMRES on (I + D^-1 V) psi = D^-1 s, left-preconditioned
//                        by (I - D^-1 V) when neumann is set (the order-1 von Neumann series
//                        of multiscat/multiscat/_scipy_von_neumann.py:40-58).
//   Mixed precision      block Jacobi with FP32 inner GMRES on the FP64 preconditioned
//                        residual r = M D^-1 (s - A psi); psi accumulates in FP64.
// Every vector and matrix lives in HBM; the host only sequences launches and reads one
// integer (instances still iterating) per Arnoldi step.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <functional>
#include <sstream>
#include <memory>
#include <cstdlib>
#include <fstream>
#include <list>
#include <map>
#include <mutex>
#include <set>

#include "../../include/multiscat_amd.h"
#include "msx_dft.h"
#include "msx_krylov.h"
#include "msx_profile.h"

using namespace msx;

namespace msx {

thread_local std::string g_last_error;
void set_last_error(const std::string &msg) { g_last_error = msg; }

template <typename T>
using Op = std::function<void(const cx<T> *in, cx<T> *out, const int *active)>;
// The operator applied to in = raw / norm[b] (per instance), also storing that scaled input to
// `in`: the normalisation of a new Krylov vector folded into the operator's first pass.
template <typename T>
using OpScaled = std::function<void(const cx<T> *raw, const double *norm, cx<T> *in, cx<T> *out,
                                    const int *active)>;

// ---------------------------------------------------------------------------------------
// Krylov workspace for one precision.
// ---------------------------------------------------------------------------------------
template <typename T>
struct Krylov {
    int n_inst = 0, kmax = 0;
    long long total = 0;
    hipStream_t s = 0;
    KLayout K{};
    KState S{};
    DBuf<long long> start_d, len_d;
    DBuf<int> enabled, kconv, iters, cycle_iters, converged;
    DBuf<int> reorth;  // instances taking the second Gram-Schmidt update (selective form)
    // active flags and the count of iterating instances in one buffer, so the host reads
    // both with a single copy
    DBuf<int> flags;
    struct IntView {
        int *p = nullptr;
    } active, count;
    // chunk partials of the projections; partial2 holds the second CGS2 pass's while the
    // first pass's are still being summed by their consumer
    DBuf<double> beta, beta0, resid, hnorm, partial, partial2;
    DBuf<cx<double>> H, g, cs, sn, y, hcol, hcol2;
    std::vector<DBuf<cx<T>>> blocks;
    std::vector<cx<T> *> vp;
    DBuf<cx<T> *> vp_d;
    DBuf<cx<T>> w;
    int *flags_h = nullptr;  // pinned, two slots of (active flags, count) in flight
    int *flags_dev = nullptr;  // the same memory as the device addresses it
    hipEvent_t ev[2] = {nullptr, nullptr};
    int next_slot = 0;
    int *mask_h = nullptr;  // pinned, two slots of n_inst flags staged for upload
    double *dbl_h = nullptr, *dbl_dev = nullptr;  // pinned n_inst doubles, and its device view
    hipEvent_t mask_ev[2] = {nullptr, nullptr};
    int mask_slot = 0;
    std::vector<long long> len_h;
    std::shared_ptr<ActiveSnap> snap;  // host view of the mask the launches issued now see
    int last_count = -1;  // iterating instances at the last count read (an upper bound later)
    // Estimate of the instances the launches issued now will see (launch shapes only): the
    // last count read is two iterations old when the next iteration is queued, and at the end
    // of a solve the count falls by a similar factor each iteration, so the estimate carries
    // the last fall forward once.
    int hint_count = -1;
    // Arnoldi iterations captured as HIP graphs, keyed by iteration, operator form, flag slot
    // and launch shape, in one set per signature (the operator's buffers and options): the
    // GRAPH_SETS signatures used last keep theirs, so solves that alternate (a forward solve
    // and its adjoint on the same workspace) replay instead of capturing again
    static constexpr int GRAPH_SETS = 4;
    std::list<std::pair<std::string, std::map<long long, hipGraphExec_t>>> graph_sets;  // newest first
    static void destroy_set(std::map<long long, hipGraphExec_t> &g) {
        for (auto &kv : g) (void)hipGraphExecDestroy(kv.second);
        g.clear();
    }
    std::map<long long, hipGraphExec_t> &graphs_for(const std::string &sig) {
        for (auto it = graph_sets.begin(); it != graph_sets.end(); ++it)
            if (it->first == sig) {
                graph_sets.splice(graph_sets.begin(), graph_sets, it);
                return graph_sets.front().second;
            }
        if ((int)graph_sets.size() >= GRAPH_SETS) {
            destroy_set(graph_sets.back().second);
            graph_sets.pop_back();
        }
        graph_sets.emplace_front(sig, std::map<long long, hipGraphExec_t>());
        return graph_sets.front().second;
    }
    void clear_graphs() {
        for (auto &g : graph_sets) destroy_set(g.second);
        graph_sets.clear();
    }

    Krylov(const std::vector<long long> &start, const std::vector<long long> &len, int kmax_,
           hipStream_t stream, long long chunk = 2048)
        : n_inst((int)start.size()), kmax(kmax_), s(stream), len_h(len) {
        if (kmax > KR_MAX_BASIS - 1) kmax = KR_MAX_BASIS - 1;
        long long maxlen = 1;
        for (int b = 0; b < n_inst; ++b) {
            total = std::max(total, start[b] + len[b]);
            maxlen = std::max(maxlen, len[b]);
        }
        K.n_inst = n_inst;
        K.kmax = kmax;
        K.chunk = chunk > 0 ? chunk : 2048;
        K.nchunk = (int)((maxlen + K.chunk - 1) / K.chunk);
        start_d.alloc(n_inst);
        len_d.alloc(n_inst);
        start_d.upload(start.data(), n_inst, s);
        len_d.upload(len.data(), n_inst, s);
        K.start = start_d.p;
        K.len = len_d.p;
        for (auto *b : {&enabled, &kconv, &iters, &cycle_iters, &converged, &reorth}) b->alloc(n_inst);
        flags.alloc((size_t)n_inst + 1);
        active.p = flags.p;
        count.p = flags.p + n_inst;
        for (auto *b : {&beta, &beta0, &resid, &hnorm}) b->alloc(n_inst);
        const size_t slots = (size_t)n_inst * (kmax + 1);
        partial.alloc(slots * K.nchunk * 2);
        partial2.alloc(slots * K.nchunk * 2);
        H.alloc(slots * (kmax + 1));
        for (auto *b : {&g, &cs, &sn, &y, &hcol, &hcol2}) b->alloc(slots);
        vp_d.alloc(kmax + 1);
        w.alloc(total);
        K.active = active.p;
        S = KState{H.p, g.p, cs.p, sn.p, y.p, beta.p, beta0.p, resid.p, hnorm.p,
                   kconv.p, iters.p, cycle_iters.p, converged.p, 0.0, 3};
        MSX_HIP(hipHostMalloc((void **)&flags_h, 2 * sizeof(int) * (n_inst + 1), hipHostMallocDefault));
        MSX_HIP(hipHostGetDevicePointer((void **)&flags_dev, flags_h, 0));
        MSX_HIP(hipHostMalloc((void **)&mask_h, 2 * sizeof(int) * n_inst, hipHostMallocDefault));
        MSX_HIP(hipHostMalloc((void **)&dbl_h, sizeof(double) * n_inst, hipHostMallocDefault));
        MSX_HIP(hipHostGetDevicePointer((void **)&dbl_dev, dbl_h, 0));
        for (auto &e : ev) MSX_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        for (auto &e : mask_ev) MSX_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        set_all_active(nullptr);
    }
    ~Krylov() {
        clear_graphs();
        if (flags_h) (void)hipHostFree(flags_h);
        if (mask_h) (void)hipHostFree(mask_h);
        if (dbl_h) (void)hipHostFree(dbl_h);
        for (auto *arr : {ev, mask_ev})
            for (int i = 0; i < 2; ++i)
                if (arr[i]) (void)hipEventDestroy(arr[i]);
    }
    // active[b] = mask ? mask[b] : 1 (mask is a host array), staged through a pinned slot and
    // uploaded in stream order without waiting for the device
    void set_all_active(const int *mask) {
        const int sl = mask_slot;
        mask_slot ^= 1;
        MSX_HIP(hipEventSynchronize(mask_ev[sl]));  // the upload that last read this slot is done
        int *h = mask_h + (size_t)sl * n_inst;
        for (int b = 0; b < n_inst; ++b) h[b] = mask ? (mask[b] ? 1 : 0) : 1;
        MSX_HIP(hipMemcpyAsync(active.p, h, sizeof(int) * n_inst, hipMemcpyHostToDevice, s));
        MSX_HIP(hipEventRecord(mask_ev[sl], s));
        snap = std::make_shared<ActiveSnap>();
        snap->on.assign(h, h + n_inst);
        last_count = 0;
        for (int b = 0; b < n_inst; ++b) last_count += h[b];
        hint_count = last_count;
    }
    void ensure(int nvec) {
        if (nvec > kmax + 1) nvec = kmax + 1;
        if ((int)vp.size() >= nvec) return;
        while ((int)vp.size() < nvec) {
            const int add = std::min(32, kmax + 1 - (int)vp.size());
            blocks.emplace_back((size_t)add * total);
            for (int i = 0; i < add; ++i) vp.push_back(blocks.back().p + (size_t)i * total);
        }
        vp_d.upload(vp.data(), vp.size(), s);
    }
    // The number of iterating instances and their mask, copied to pinned slot `slot` behind
    // the work queued so far; launches issued after this see a mask not known yet (a new
    // snapshot, filled by wait()).
    struct Pending {
        int slot = -1;
        std::shared_ptr<ActiveSnap> next;
    };
    // Device address of the slot the next post() takes: the kernel that ends an iteration
    // writes the flags and the count there itself (post(true) then only records the event).
    // Its previous contents were read by the wait() one iteration earlier.
    int *next_flags() const { return flags_dev + (size_t)next_slot * (n_inst + 1); }
    Pending post(bool written = false) {
        const int slot = next_slot;
        next_slot ^= 1;
        if (!written)
            MSX_HIP(hipMemcpyAsync(flags_h + (size_t)slot * (n_inst + 1), flags.p,
                                   sizeof(int) * (n_inst + 1), hipMemcpyDeviceToHost, s));
        MSX_HIP(hipEventRecord(ev[slot], s));
        Pending p{slot, std::make_shared<ActiveSnap>()};
        p.next->ready.store(false, std::memory_order_relaxed);
        snap = p.next;
        return p;
    }
    int wait(const Pending &p) {
        MSX_HIP(hipEventSynchronize(ev[p.slot]));
        const int *a = flags_h + (size_t)p.slot * (n_inst + 1);
        p.next->on.assign(a, a + n_inst);
        p.next->ready.store(true, std::memory_order_release);
        const int before = last_count;
        last_count = a[n_inst];
        hint_count = before > last_count && last_count > 0
                         ? std::max(1, (int)((long long)last_count * last_count / before))
                         : last_count;
        return last_count;
    }
    int read_count() { return wait(post()); }
    // per-instance work of a launch over the iterating instances (counted against snap)
    Work per_instance(double per_elem) const {
        Work w;
        if (!profiling()) return w;
        w.per.resize(n_inst);
        for (int b = 0; b < n_inst; ++b) w.per[b] = per_elem * (double)len_h[b];
        w.snap = snap;
        return w;
    }
    std::vector<int> host_ints(const DBuf<int> &b) {
        std::vector<int> h(n_inst);
        b.download(h.data(), n_inst, s);
        MSX_HIP(hipStreamSynchronize(s));
        return h;
    }
    std::vector<double> host_doubles(const DBuf<double> &b) {
        std::vector<double> h(n_inst);
        b.download(h.data(), n_inst, s);
        MSX_HIP(hipStreamSynchronize(s));
        return h;
    }
    // Per-instance squared norms of an arbitrary vector with this layout (host result), written
    // by the device straight into pinned host memory (no staged copy of the whole column
    // array, n_inst (kmax + 1) values, to read one per instance)
    std::vector<double> norms2(const cx<T> *v) {
        k_dots<T>(K, vp_d.p, 0, v, 1, partial.p, hcol.p, 0, s);
        k_gather_re(n_inst, kmax + 1, hcol.p, dbl_dev, s);
        MSX_HIP(hipStreamSynchronize(s));
        return std::vector<double>(dbl_h, dbl_h + n_inst);
    }
};

// GMRES(kmax) on op x = rhs for the instances active on entry. x accumulates the Krylov
// corrections (x += V y each cycle). If x_zero, the first residual is rhs itself (x0 = 0,
// as in the reference); otherwise r0 = rhs - op(x) (needs TX == T). Restarts continue from
// the current x until maxiter iterations or convergence. basis_order: the vector order of the
// basis when it differs from x's (node-major FP32 basis, channel-major FP64 x). eps_inst
// (device, optional): a tolerance per instance instead of eps. op_scaled (optional): the
// operator with the normalisation of its input folded in; each new basis vector is then
// normalised by the next operator call instead of a separate pass.
// graph_sig (optional): replay each Arnoldi iteration from a HIP graph captured the first
// time that (iteration, operator form, flag slot, launch shape) comes up; graph_sig names the
// operator's buffers and options, and selects the set of captured graphs (the workspace keeps
// the sets of its last few signatures). shape_key
// gives the host-side launch choices that depend on the active count (part of the key).
// stop_below: once the host sees at most this many instances iterating, the solve stops after
// the iteration already queued and every instance keeps the solution of its current subspace.
// Profiling needs the eager launches (its events bracket single kernels), so it turns this off.
// orth_passes: Gram-Schmidt passes (1, 2), or 3: one pass and the second only for the
// instances whose first left the new vector's projection on the basis above 1e-4 (up to
// k + 1 = 16; two passes above). Returns the number of iterations queued.
template <typename T, typename TX>
int gmres(Krylov<T> &kr, const Op<T> &op, const cx<T> *rhs, cx<TX> *x, bool x_zero,
           double eps, int window, int maxiter, bool allow_restart, int orth_passes = 2,
           const Layout *basis_order = nullptr, const double *eps_inst = nullptr,
           const OpScaled<T> *op_scaled = nullptr, const std::string *graph_sig = nullptr,
           const std::function<int()> *shape_key = nullptr, int stop_below = 0) {
    hipStream_t s = kr.s;
    const bool use_graphs = graph_sig != nullptr && !profiling();
    std::map<long long, hipGraphExec_t> *gset = nullptr;  // this signature's captured iterations
    if (use_graphs) {
        // the recurrence state goes to the Givens kernel by value: its tolerance, window and
        // per-instance tolerances are part of what a captured iteration replays
        char extra[160];
        std::snprintf(extra, sizeof extra, "|eps%.17g|win%d|einst%p|orth%d|h%g,%g", eps, window,
                      (const void *)eps_inst, orth_passes, kr.S.hscale, kr.S.hshift);
        gset = &kr.graphs_for(*graph_sig + extra);
        // every flag slot posted by an earlier solve has been waited for: start at slot 0, so
        // iteration k of a cycle always writes the same slot
        kr.next_slot = 0;
    }
    kr.S.eps = eps;
    kr.S.eps_inst = eps_inst;
    kr.S.window = window;
    k_reset(kr.K, kr.S, kr.enabled.p, s);
    kr.ensure(1);
    int done = 0;
    bool first = true;
    for (;;) {
        cx<T> *v0 = kr.vp[0];
        // the first residual r0: rhs itself for x0 = 0 (scaled straight into v0 below)
        const cx<T> *r0 = rhs;
        if (!(x_zero && first)) {
            if constexpr (std::is_same<T, TX>::value) {
                op(reinterpret_cast<const cx<T> *>(x), kr.w.p, kr.active.p);
                k_axpby<T>(kr.K, rhs, kr.w.p, T(-1), v0, kr.active.p, s);
                r0 = v0;
            } else {
                throw Error(3, "restart of a mixed-precision inner solve");
            }
        }
        k_dots<T>(kr.K, kr.vp_d.p, 0, r0, 1, kr.partial.p, kr.hcol.p, 0, s);
        // with an initial guess the tolerance stays relative to |rhs| (not to the first
        // residual), so a good guess saves iterations
        const bool ref_rhs = first && !x_zero;
        if (ref_rhs) k_dots<T>(kr.K, kr.vp_d.p, 0, rhs, 1, kr.partial.p, kr.hcol2.p, 0, s);
        k_cycle_start(kr.K, kr.S, kr.hcol.p, first ? 1 : 0, kr.count.p, s,
                      ref_rhs ? kr.hcol2.p : nullptr, kr.next_flags());
        k_scale<T>(kr.K, r0, v0, kr.beta.p, kr.active.p, s);
        // The host learns whether any instance still iterates one iteration late: iteration k
        // is queued before the count of iteration k - 1 is read, so the GPU never waits for
        // the host. Every kernel skips inactive instances, so an iteration queued after the
        // last instance converged does nothing.
        int alive = 1;
        typename Krylov<T>::Pending prev = kr.post(true);  // the count after the cycle start
        bool pending_norm = false;  // vp[k] is still kr.w (unnormalised, norm in kr.hnorm)
        for (int k = 0; k < kr.kmax && done < maxiter; ++k) {
            if (prev.slot >= 0 && k + 2 > (int)kr.vp.size()) {
                // settle before allocating another block of basis vectors
                alive = kr.wait(prev);
                prev.slot = -1;
                if (alive == 0) break;
            }
            kr.ensure(k + 2);
            const bool scaled_in = pending_norm;
            // the fixed kernel sequence of iteration k: operator, Gram-Schmidt passes, Givens
            // update (writing the flags to slot next_slot), normalisation
            auto iteration = [&]() {
            const double *norm_src = kr.partial.p;  // where the Givens kernel finds |w|^2 partials
            if (scaled_in) {
                (*op_scaled)(kr.w.p, kr.hnorm.p, kr.vp[k], kr.w.p, kr.active.p);
            } else {
                op(kr.vp[k], kr.w.p, kr.active.p);
            }
            {
                // CGS2: two passes of (k+1 dot products, then w -= V h); HBM traffic of the
                // vector kernels is 4 (k+1) + 6 vectors of the iterating instances, 3 (k+1) + 6
                // when the first update and the second projection share a pass (k+1 <= 16);
                // one pass: 2 (k+1) + 3; the selective form (3): 2 (k+1) + 3 plus the second
                // update's k + 3 where an instance takes it (counted here as if none did)
                const bool selective = orth_passes == 3 && k + 1 <= 16;
                const double P = orth_passes >= 2 && !selective ? 2.0 : 1.0;
                const double vecs = selective ? 2.0 * (k + 1) + 3.0
                                    : orth_passes >= 2 ? (k + 1 <= 16 ? 3.0 * (k + 1) + 6.0 : 4.0 * (k + 1) + 6.0)
                                                       : 2.0 * (k + 1) + 3.0;
                const Work fl = kr.per_instance(P * 16.0 * (k + 1));
                const Work by = kr.per_instance(vecs * sizeof(cx<T>));
                const char *nm = selective ? (sizeof(T) == 4 ? "arnoldi_cgs1s_f32" : "arnoldi_cgs1s_f64")
                                 : orth_passes >= 2
                                     ? (sizeof(T) == 4 ? "arnoldi_cgs2_f32" : "arnoldi_cgs2_f64")
                                     : (sizeof(T) == 4 ? "arnoldi_cgs1_f32" : "arnoldi_cgs1_f64");
                ProfScope ps(nm, s, fl, by);
                if (selective) {
                    // one pass, and the second only for the instances whose first left the new
                    // vector's projection on the basis above 1e-4 of its norm in FP32, 1e-8
                    // (about the square root of the unit roundoff) in FP64 (the second
                    // projection and |w|^2 come with the first update's pass over V)
                    k_dots<T>(kr.K, kr.vp_d.p, k + 1, kr.w.p, 0, kr.partial.p, nullptr, 0, s);
                    k_axpy_dots<T>(kr.K, kr.vp_d.p, k + 1, nullptr, kr.w.p, kr.partial2.p, nullptr,
                                   s, kr.partial.p, kr.hcol.p, 1);
                    k_reorth_select(kr.K, k + 1, kr.partial2.p, kr.reorth.p, kr.hcol2.p,
                                    sizeof(T) == 4 ? 1e-4 : 1e-8, s);
                    KLayout K2 = kr.K;
                    K2.active = kr.reorth.p;
                    k_axpy<T>(K2, kr.vp_d.p, k + 1, nullptr, kr.w.p, 1, kr.partial2.p, nullptr, s,
                              kr.partial2.p, kr.hcol2.p);
                    norm_src = kr.partial2.p;
                } else if (orth_passes >= 2 && k + 1 <= 16) {
                    // each pass's chunk partials summed by the kernel that consumes them (no
                    // reduction launches): dots -> partial; the first update sums them (and
                    // stores hcol) and writes the second projection's partials to partial2;
                    // the second update sums those (and stores hcol2) and adds the norm
                    // partials in slot k + 1 of partial2, which the Givens kernel sums
                    k_dots<T>(kr.K, kr.vp_d.p, k + 1, kr.w.p, 0, kr.partial.p, nullptr, 0, s);
                    k_axpy_dots<T>(kr.K, kr.vp_d.p, k + 1, nullptr, kr.w.p, kr.partial2.p, nullptr,
                                   s, kr.partial.p, kr.hcol.p);
                    k_axpy<T>(kr.K, kr.vp_d.p, k + 1, nullptr, kr.w.p, 1, kr.partial2.p, nullptr, s,
                              kr.partial2.p, kr.hcol2.p);
                    norm_src = kr.partial2.p;
                } else if (orth_passes >= 2) {
                    k_dots<T>(kr.K, kr.vp_d.p, k + 1, kr.w.p, 0, kr.partial.p, kr.hcol.p, 0, s);
                    // the first update and the second projection share one pass over V
                    if (!k_axpy_dots<T>(kr.K, kr.vp_d.p, k + 1, kr.hcol.p, kr.w.p, kr.partial.p,
                                        kr.hcol2.p, s)) {
                        k_axpy<T>(kr.K, kr.vp_d.p, k + 1, kr.hcol.p, kr.w.p, 0, kr.partial.p,
                                  nullptr, s);
                        k_dots<T>(kr.K, kr.vp_d.p, k + 1, kr.w.p, 0, kr.partial.p, kr.hcol2.p, 0, s);
                    }
                    k_axpy<T>(kr.K, kr.vp_d.p, k + 1, kr.hcol2.p, kr.w.p, 1, kr.partial.p,
                              nullptr, s);
                } else {
                    k_dots<T>(kr.K, kr.vp_d.p, k + 1, kr.w.p, 0, kr.partial.p, kr.hcol.p, 0, s);
                    k_axpy<T>(kr.K, kr.vp_d.p, k + 1, kr.hcol.p, kr.w.p, 1, kr.partial.p, nullptr,
                              s);
                }
            }
            // the Givens kernel sums the norm partials of the last update itself
            k_givens(kr.K, k, kr.hcol.p, orth_passes >= 2 ? kr.hcol2.p : nullptr, kr.S, kr.count.p, s,
                     kr.next_flags(), norm_src);
            // v_{k+1} = w / h_{k+1,k}: by the next operator call when it can, else here (a
            // vector the loop leaves unnormalised is never read: the update uses v_0 .. v_k)
            if (!op_scaled) k_scale<T>(kr.K, kr.w.p, kr.vp[k + 1], kr.hnorm.p, kr.active.p, s);
            };
            if (use_graphs) {
                const long long shape = shape_key ? (long long)(*shape_key)() : 0;
                const long long key =
                    ((((long long)k * 2 + (scaled_in ? 1 : 0)) * 2 + kr.next_slot) << 20) + shape;
                auto it = gset->find(key);
                hipGraphExec_t exec = nullptr;
                if (it == gset->end()) {
                    hipGraph_t graph = nullptr;
                    MSX_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
                    try {
                        iteration();
                    } catch (...) {
                        hipGraph_t partial_graph = nullptr;
                        (void)hipStreamEndCapture(s, &partial_graph);
                        if (partial_graph) (void)hipGraphDestroy(partial_graph);
                        throw;
                    }
                    MSX_HIP(hipStreamEndCapture(s, &graph));
                    const hipError_t e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
                    (void)hipGraphDestroy(graph);
                    MSX_HIP(e);
                    (*gset)[key] = exec;
                } else {
                    exec = it->second;
                }
                MSX_HIP(hipGraphLaunch(exec, s));
            } else {
                iteration();
            }
            pending_norm = op_scaled != nullptr;
            ++done;
            const typename Krylov<T>::Pending cur = kr.post(true);
            if (prev.slot >= 0) {
                alive = kr.wait(prev);
                if (alive == 0) {  // iteration k had nothing left to do
                    --done;
                    alive = kr.wait(cur);
                    prev.slot = -1;
                    break;
                }
                if (alive <= stop_below) {  // the few still iterating stop after iteration k
                    prev = cur;
                    break;
                }
            }
            prev = cur;
        }
        if (prev.slot >= 0) alive = kr.wait(prev);
        k_backsolve(kr.K, kr.S, nullptr, s);
        if (basis_order && basis_order->nm) {
            if constexpr (std::is_same<TX, double>::value)
                k_update_nm<T>(*basis_order, kr.kmax, kr.vp_d.p, kr.cycle_iters.p, kr.y.p, x, s);
            else
                throw Error(3, "node-major basis: the solution must be FP64");
        } else {
            k_update<T, TX>(kr.K, kr.vp_d.p, kr.cycle_iters.p, kr.y.p, x, nullptr, s);
        }
        first = false;
        x_zero = false;
        if (alive == 0 || done >= maxiter || !allow_restart) break;
        k_reactivate(kr.K, kr.S, kr.enabled.p, s);
    }
    return done;
}

}  // namespace msx

// ---------------------------------------------------------------------------------------
// Potential table.
// ---------------------------------------------------------------------------------------
// Eigensystem of the symmetric n x n matrix A (destroyed; Vm n x n workspace): method 0
// divide and conquer when n fits it, else QL; 1 QL; 2 divide and conquer.
static void run_eigensolver(int n, double *A, double *Vm, DBuf<double> &dcw, double *lam,
                            double *Q, double *QT, int *info, long long *ticks, hipStream_t s,
                            int method) {
    const bool dc = method == 2 || (method == 0 && eig_dc_supported(n));
    const double nn = n;
    if (dc) {
        if (!eig_dc_supported(n)) throw Error(MSX_ERR_INVALID, "divide and conquer: 2 <= n <= 1024");
        if (dcw.n < eig_dc_workspace(n)) dcw.alloc(eig_dc_workspace(n));
        // tridiagonalisation (4/3 n^3), merge products (about 8/3 n^3 over the tree levels),
        // back-transformation by the reflectors (2 n^3)
        ProfScope ps("eigensystem_f64", s, (4.0 / 3.0 + 8.0 / 3.0 + 2.0) * nn * nn * nn,
                     5.0 * 8.0 * nn * nn);
        launch_eigensystem_dc(n, A, dcw.p, lam, Q, QT, info, ticks, s);
    } else {
        // tridiagonalisation and Q_H (8/3 n^3), QL rotations on Z (about 6 n^3 for about 1.5
        // sweeps per eigenvalue)
        ProfScope ps("eigensystem_f64", s, (8.0 / 3.0 + 6.0) * nn * nn * nn, 5.0 * 8.0 * nn * nn);
        launch_eigensystem(n, A, Vm, lam, Q, QT, info, ticks, s);
    }
}

struct msx_table {
    int nkx = 0, nky = 0, nz = 0;
    hipStream_t stream = 0;
    DBuf<cx<double>> V, Vt;
    DBuf<cx<float>> Vf, Vtf;
    DBuf<double> T, A, Vm, lam, Q, QT, Qp, QTp, dcw;
    DBuf<double> H0, H0p;  // symmetrised H0 and its zero-padded copy (nzp x nzp)
    DBuf<float> Qf, QTf, Qpf, QTpf;
    DBuf<_Float16> H3;  // split-FP16 planes of Q^T and Q (nz <= 224)
    int nq3 = 0;
    DBuf<int> info;
    DBuf<long long> ticks;  // phase stamps of the last eigensolve
    int nzp = 0;          // nz rounded up to a multiple of 16 (matrix-core tiles)
    DftStore dft, dft_t;  // convolution form of the coupling (table and flipped table)
    bool flipped = false;
    ~msx_table() {
        if (stream) (void)hipStreamDestroy(stream);
    }
    size_t nG() const { return (size_t)nkx * nky; }
    EigTables<double> eig64() const { return {nzp, Qp.p, QTp.p, lam.p}; }
    EigTables<float> eig32() const { return {nzp, Qpf.p, QTpf.p, lam.p}; }
    H3Planes h3() const {
        H3Planes h;
        if (nq3 > 0) h = H3Planes{H3.p, nq3, Qpf.p + (size_t)(nz - 1) * nzp};
        return h;
    }
    void prepare() {
        const size_t n2 = (size_t)nz * nz;
        launch_build_h0(nz, T.p, V.p, A.p, stream);
        run_eigensolver(nz, A.p, Vm.p, dcw, lam.p, Q.p, QT.p, info.p, ticks.p, stream, 0);
        int h_info = 0;
        info.download(&h_info, 1, stream);
        MSX_HIP(hipStreamSynchronize(stream));
        if (h_info != 0) throw Error(MSX_ERR_EIGEN, "eigensolver did not converge");
        launch_to_float_real((long long)n2, Q.p, Qf.p, stream);
        launch_to_float_real((long long)n2, QT.p, QTf.p, stream);
        nzp = (nz + 15) / 16 * 16;
        const size_t p2 = (size_t)nzp * nzp;
        if (Qp.n != p2) {
            Qp.alloc(p2);
            QTp.alloc(p2);
            Qpf.alloc(p2);
            QTpf.alloc(p2);
        }
        launch_pad_square(nz, nzp, Q.p, Qp.p, stream);
        launch_pad_square(nz, nzp, QT.p, QTp.p, stream);
        launch_pad_square(nz, nzp, Q.p, Qpf.p, stream);
        launch_pad_square(nz, nzp, QT.p, QTpf.p, stream);
        nq3 = h3_padded_size(nz);
        if (nq3 > 0) {
            if (H3.n != h3_plane_halves(nq3)) H3.alloc(h3_plane_halves(nq3));
            launch_h3_planes(nz, nq3, Q.p, H3.p, stream);
        }
        if (H0p.n != p2) {
            H0.alloc(n2);
            H0p.alloc(p2);
        }
        launch_build_h0(nz, T.p, V.p, H0.p, stream);
        launch_pad_square(nz, nzp, H0.p, H0p.p, stream);
        launch_to_float((long long)(nG() * nz), V.p, Vf.p, stream);
        build_dft_tables(nkx, nky, nz, V.p, dft, stream);
        flipped = false;
    }
    void ensure_flipped() {
        if (flipped) return;
        if (!Vt.p) {
            Vt.alloc(nG() * nz);
            Vtf.alloc(nG() * nz);
        }
        launch_flip_table(nkx, nky, nz, V.p, Vt.p, stream);
        launch_to_float((long long)(nG() * nz), Vt.p, Vtf.p, stream);
        build_dft_tables(nkx, nky, nz, Vt.p, dft_t, stream);
        flipped = true;
    }
};

// ---------------------------------------------------------------------------------------
// Batch of conditions.
// ---------------------------------------------------------------------------------------
struct msx_batch {
    msx_table *t = nullptr;
    int n_inst = 0, max_nc = 0;
    long long total_ch = 0;
    std::vector<int> nc, spec;
    std::vector<long long> off;
    DBuf<int> nc_d, chx_d, chy_d, spec_d, chmap_d, iters_acc;
    DBuf<double> eps_in_d;  // per-instance tolerances of the mixed solver's inner solves
    DBuf<long long> off_d;
    DBuf<double> e_d, fac_d, P_d, dLdP_d;
    DBuf<float> facf_d, facf_nm_d;  // FP32 factors, channel-major and node-major
    DBuf<cx<double>> c_d, b_d, a0_d, bsrc_d, psi_d, lam_d;
    DBuf<cx<float>> cf_d;
    // work vectors and Krylov workspaces (allocated on first use, reused by later solves)
    DBuf<cx<double>> source, src, rhs, tmp1, tmp2, tmp3;
    DBuf<cx<double>> adj_src, adj_state, grad_d, grad_part;  // adjoint solve and gradient
    DBuf<cx<float>> f_rhs, f_tmp1, f_tmp2;
    std::unique_ptr<Krylov<double>> kr64, kr_outer;
    std::unique_ptr<Krylov<float>> kr32;
    std::vector<int> iters, info;
    std::vector<double> resid;
    bool factors_ready = false, facf_nm_ready = false, solved = false;
    // psi_d holds the node-major order (the mixed solver's refinement runs in that order and
    // leaves the state so); psi_channel_major() restores the batch layout for consumers
    bool psi_nm = false;
    struct KdimMemo {
        int restart = -1, maxiter = -1, dim = 0;
    } kdim_memo[2];  // Krylov dimension per precision (FP64, mixed) for the last options seen
    // Launch shapes chosen by the autotuner (built-in values until it runs)
    struct Tuning {
        int dft_npw32 = 0, dft_npw64 = 0;  // DFT coupling: nodes per wavefront (0: built-in 4)
        long long chunk32 = 2048, chunk64 = 2048;  // Arnoldi vector kernels: elements per block
        int diag64_cpw = 0;                // FP64 diagonal blocks: channels per wavefront
        int diag32_cpw = 0;                // split-FP16 FP32 D^-1: channels per wavefront (0: 16)
        bool done = false;
        std::string key, report;
    } tuning;
    // Arnoldi block lengths chosen at the batch's first tuned solve of each precision (0: none
    // yet); later keys of the batch (another vector order: a forward solve and its adjoint)
    // keep them, so the Krylov workspaces sized for them are not rebuilt
    long long kept_chunk32 = 0, kept_chunk64 = 0;

    long long n_elem() const { return total_ch * t->nz; }
    Layout layout() const {
        return Layout{t->nkx, t->nky, t->nz, n_inst, max_nc, nc_d.p, off_d.p, chx_d.p, chy_d.p,
                      (long long)(n_elem() / t->nz)};
    }
    std::vector<long long> kstart() const {
        std::vector<long long> v(n_inst);
        for (int b = 0; b < n_inst; ++b) v[b] = off[b] * t->nz;
        return v;
    }
    std::vector<long long> klen() const {
        std::vector<long long> v(n_inst);
        for (int b = 0; b < n_inst; ++b) v[b] = (long long)nc[b] * t->nz;
        return v;
    }
    template <typename T>
    static void need(DBuf<T> &b, size_t n) {
        if (b.n < n) b.alloc(n);
    }
    // the state in the given order (through tmp2, which holds nothing between calls)
    void psi_order(bool nm) {
        if (psi_nm == nm) return;
        const size_t n = (size_t)n_elem();
        need(tmp2, n);
        launch_state_order(layout(), psi_d.p, tmp2.p, nm, t->stream);
        MSX_HIP(hipMemcpyAsync(psi_d.p, tmp2.p, sizeof(cx<double>) * n, hipMemcpyDeviceToDevice,
                               t->stream));
        psi_nm = nm;
    }
    void psi_channel_major() { psi_order(false); }
    void build_factors() {
        ProfScope ps("factors_f64", t->stream, 2.0 * total_ch * t->nz * t->nz,
                     (double)total_ch * t->nz * 8.0);
        if (!launch_factors_mfma(total_ch, t->nz, t->eig64(), e_d.p, fac_d.p, t->stream))
            launch_preconditioner_factors_qt(total_ch, t->nz, t->Q.p, t->QT.p, t->lam.p, e_d.p,
                                             fac_d.p, t->stream);
        launch_to_float_real(total_ch * t->nz, fac_d.p, facf_d.p, t->stream);
        factors_ready = true;
        facf_nm_ready = false;
    }
};

namespace {

using namespace msx;

// ---- operators -------------------------------------------------------------------------
struct Ops {
    msx_batch &B;
    msx_table &t;
    Layout L;
    Layout L32;  // order of the FP32 vectors (set_fp32_order)
    Layout L64;  // order of the FP64 vectors the full coupling takes (L unless node-major)
    bool neumann;
    int order = 1;  // Jacobi: degree of the Neumann factor M = sum_{i<=order} (-D^-1 V)^i
    const cx<double> *V;  // coupling table (flipped for the transposed system)
    const cx<float> *Vf;
    const DftStore *dft;  // its convolution form
    bool direct;          // force the direct sum for the full coupling
    bool vector_diag = false;  // force the vector-ALU D^-1 kernel
    bool split_f16 = true;            // split-FP16 FP32 D^-1 where the kernel covers the shape
    bool split_f16_required = false;  // ... and an error where it does not
    hipStream_t s;

    Ops(msx_batch &b, bool neu, bool transposed, bool force_direct)
        : B(b), t(*b.t), L(b.layout()), neumann(neu), s(b.t->stream) {
        if (transposed) {
            t.ensure_flipped();
            V = t.Vt.p;
            Vf = t.Vtf.p;
            dft = &t.dft_t;
        } else {
            V = t.V.p;
            Vf = t.Vf.p;
            dft = &t.dft;
        }
        direct = force_direct || dft->np == 0;
        L32 = L;
        L64 = L;
    }
    // FP32 vectors node-major when every FP32 kernel of this configuration reads and writes
    // that order: the DFT coupling of tables up to 32 x 32 and the diagonal blocks with Q^T in
    // LDS (nzp <= 160) or the split-FP16 ones with their planes in L2 (nz <= 224)
    void set_fp32_order() {
        L32 = L;
        bool nm_diag = t.nzp > 0 && t.nzp <= 160;
        if (!nm_diag && split_f16 && t.nzp <= 224) {
            Layout Ln = L;
            Ln.nm = 1;
            nm_diag = diag_h3_fits(Ln, t.h3());
        }
        L32.nm = (!direct && !vector_diag && nm_diag && dft->np <= 32) ? 1 : 0;
        if (L32.nm && !B.facf_nm_ready) {  // the FP32 factors in the same order as the vectors
            msx_batch::need(B.facf_nm_d, (size_t)B.n_elem());
            launch_to_float_real_nm(L32, B.fac_d.p, B.facf_nm_d.p, s);
            B.facf_nm_ready = true;
        }
    }
    template <typename T>
    const Layout &lay() const {
        return sizeof(T) == 4 ? L32 : L;
    }
    // full off-diagonal coupling y = V x (DFT form on the matrix cores when it applies)
    // ---- algorithmic work of one launch over the instances it processes (profiling) ----
    // host view of the Krylov mask the launches see, if any, and its estimate of the count
    const std::shared_ptr<ActiveSnap> *snap_src = nullptr;
    const int *count_src = nullptr;
    int active_hint(const int *act) const { return act && count_src ? *count_src : -1; }
    // The launch choices that depend on the host's count of iterating instances (the DFT
    // coupling's nodes per workgroup, FP32 and FP64): part of a captured iteration's key.
    int launch_shape_key() const {
        if (direct || !count_src) return 0;
        const int hint = *count_src;
        return dft_block_nodes(L32, false, hint, B.tuning.dft_npw32) |
               (dft_block_nodes(L, true, hint, B.tuning.dft_npw64) << 8);
    }
    // Everything a captured operator launch depends on besides the iteration: buffers, tables,
    // options and launch shapes.
    std::string graph_signature(const char *what, const void *a, const void *b,
                                const void *c) const {
        char buf[512];
        std::snprintf(buf, sizeof buf,
                      "%s|%p|%p|%p|V%p|Vf%p|D%p|nm%d|neu%d|ord%d|dir%d|vd%d|h3%d|npw%d,%d|cpw%d,%d|fac%p",
                      what, a, b, c, (const void *)V, (const void *)Vf, (const void *)dft,
                      L32.nm, neumann ? 1 : 0, order, direct ? 1 : 0, vector_diag ? 1 : 0,
                      split_f16 ? 1 : 0,
                      B.tuning.dft_npw32, B.tuning.dft_npw64, B.tuning.diag64_cpw, B.tuning.diag32_cpw,
                      (const void *)(L32.nm ? B.facf_nm_d.p : B.facf_d.p));
        return buf;
    }
    template <typename F>
    Work over_active(const int *act, F f) const {
        Work w;
        if (!profiling()) return w;
        w.per.resize(B.n_inst);
        for (int b = 0; b < B.n_inst; ++b) w.per[b] = f(B.nc[b]);
        if (act && snap_src && *snap_src) w.snap = *snap_src;
        return w;
    }
    // full coupling: direct sum 8 nc (nc-1) nz flops; DFT form 2 complex 2-D transforms of
    // the grid per node (matrix products, 8 flops per complex MAC) plus the pointwise product
    void couple_full_work(const int *act, size_t s_cx, bool dft_form, Work &fl, Work &by,
                          bool with_copy = false) {
        const double nz = L.nz, nkx = L.nkx, nky = L.nky, nG = nkx * nky;
        if (dft_form)
            fl = over_active(act, [&](int) { return nz * (16.0 * nG * (nkx + nky) + 6.0 * nG); });
        else
            fl = over_active(act, [&](int nc) { return 8.0 * nc * (nc - 1.0) * nz; });
        const double vecs = with_copy ? 3.0 : 2.0;  // in, out (and the scaled copy of in)
        by = over_active(act, [&](int nc) { return vecs * nc * nz * s_cx; }) + nG * nz * s_cx;
    }
    // in_norm / in_copy (DFT form only): couple in / in_norm[b] and store that to in_copy
    void couple_full(const cx<double> *in, cx<double> *out, const int *act,
                     const double *in_norm = nullptr, cx<double> *in_copy = nullptr) {
        Work fl, by;
        if (profiling()) couple_full_work(act, sizeof(cx<double>), !direct, fl, by, in_copy != nullptr);
        ProfScope ps(direct ? "coupling_direct_f64" : "coupling_dft_f64", s, fl, by);
        if (direct) launch_coupling<double>(L64, V, in, out, COUPLE_FULL, act, s);
        else launch_dft_coupling<double>(L64, dft->d64(), in, out, act, s, active_hint(act), in_norm,
                                         in_copy, B.tuning.dft_npw64);
    }
    void couple_full(const cx<float> *in, cx<float> *out, const int *act,
                     const double *in_norm = nullptr, cx<float> *in_copy = nullptr) {
        Work fl, by;
        if (profiling()) couple_full_work(act, sizeof(cx<float>), !direct, fl, by, in_copy != nullptr);
        if (!direct && split_f16 && dft_h3_fits(L32, dft->d32())) {
            ProfScope ps("coupling_dft_h3", s, fl, by);
            launch_dft_coupling_h3(L32, dft->d32(), in, out, act, s, active_hint(act), in_norm,
                                   in_copy, B.tuning.dft_npw32);
            return;
        }
        ProfScope ps(direct ? "coupling_direct_f32" : "coupling_dft_f32", s, fl, by);
        if (direct) launch_coupling<float>(L32, Vf, in, out, COUPLE_FULL, act, s);
        else launch_dft_coupling<float>(L32, dft->d32(), in, out, act, s, active_hint(act), in_norm,
                                        in_copy, B.tuning.dft_npw32);
    }
    void upper(const cx<double> *in, cx<double> *out, const int *act) {
        Work fl, by;
        if (profiling()) {
            fl = over_active(act, [&](int nc) { return 4.0 * nc * (nc - 1.0) * L.nz; });
            by = over_active(act, [&](int nc) { return 32.0 * nc * L.nz; }) +
                 16.0 * L.nkx * L.nky * L.nz;
        }
        ProfScope ps("coupling_upper_f64", s, fl, by);
        launch_coupling<double>(L, V, in, out, COUPLE_UPPER, act, s);
    }
    // L^-1: the lower coupling sum (4 nc (nc-1) nz) and two real nz x nz products per channel
    void lower(cx<double> *x, const cx<double> *base, cx<double> *out, double sign,
               const int *act) {
        Work fl, by;
        if (profiling()) {
            const double nz = L.nz;
            fl = over_active(act, [&](int nc) { return 4.0 * nc * (nc - 1.0) * nz + 8.0 * nc * nz * nz; });
            by = over_active(act, [&](int nc) { return 48.0 * nc * nz; }) + 16.0 * nz * nz;
        }
        ProfScope ps("lower_solve_f64", s, fl, by);
        launch_lower_solve(L, V, t.Q.p, t.QT.p, t.lam.p, B.e_d.p, B.c_d.p, B.fac_d.p, x, base, out,
                           sign, act, s);
    }
    // Gauss-Seidel, FP64: out = M (I + L^-1 U) in
    void gs(const cx<double> *in, cx<double> *out, const int *act) {
        cx<double> *u = B.tmp1.p;
        cx<double> *w1 = neumann ? B.tmp2.p : out;
        upper(in, u, act);
        lower(u, in, w1, 1.0, act);
        if (neumann) {
            upper(w1, u, act);
            lower(u, w1, out, -1.0, act);
        }
    }
    // Gauss-Seidel right-hand side: rhs = M L^-1 src (src is overwritten with L^-1 src)
    void gs_rhs(cx<double> *src, cx<double> *rhs) {
        lower(src, nullptr, nullptr, 0.0, nullptr);
        if (neumann) {
            cx<double> *u = B.tmp1.p;
            upper(src, u, nullptr);
            lower(u, src, rhs, -1.0, nullptr);
        } else {
            MSX_HIP(hipMemcpyAsync(rhs, src, sizeof(cx<double>) * B.n_elem(),
                                   hipMemcpyDeviceToDevice, s));
        }
    }
    // Block Jacobi: out = M (I + D^-1 V) in, in FP64 or FP32
    // out = base + sign * D^-1 in (base may be null): matrix cores when nz <= 320
    template <typename T>
    void dinv(const cx<T> *in, const cx<T> *base, cx<T> *out, T sign, const int *act) {
        const bool mfma = !vector_diag && diag_mfma_supported(L.nz);
        bool h3 = false;
        if constexpr (sizeof(T) == 4) h3 = mfma && split_f16 && diag_h3_fits(lay<float>(), t.h3());
        Work fl, by;
        if (profiling()) {  // two real nz x nz products on the re and im rows of each channel
            // bytes: x, y (and base) per element, the factor vectors f_j per element except in
            // the split kernel (which forms corr_j Q^T f_j in registers), e_j, c_j and f_j(N)
            // per channel, and the eigenvector tables (the split kernel's four FP16 planes)
            const double nz = L.nz;
            const double per_elem = (base ? 3.0 : 2.0) * sizeof(cx<T>) + (h3 ? 0.0 : sizeof(T));
            const double per_ch = sizeof(double) + sizeof(cx<T>) + sizeof(T);
            const double nq = t.h3().nq;
            fl = over_active(act, [&](int nc) { return nc * (8.0 * nz * nz + 12.0 * nz); });
            by = over_active(act, [&](int nc) { return nc * (nz * per_elem + per_ch); }) +
                 (h3 ? 4.0 * nq * (nq + 8) * sizeof(_Float16) : 2.0 * nz * nz * sizeof(T));
        }
        if constexpr (sizeof(T) == 4) {
            if (h3) {
                ProfScope ps("diag_h3_f32", s, fl, by);
                launch_diag_h3(lay<float>(), t.h3(), t.lam.p, B.e_d.p, c<float>(), fac<float>(), in,
                               base, out, sign, act, s, B.tuning.diag32_cpw);
                return;
            }
            if (split_f16_required)
                throw Error(MSX_ERR_INVALID,
                            "inner_products = 2: split-FP16 diagonal blocks need node-major FP32 "
                            "vectors, nz <= 224 and LDS for the batch's tile scan");
        }
        const char *name = sizeof(T) == 4 ? (mfma ? "diag_mfma_f32" : "diag_vector_f32")
                                          : (mfma ? "diag_mfma_f64" : "diag_vector_f64");
        ProfScope ps(name, s, fl, by);
        if (mfma)
            launch_diag_blocks<T>(lay<T>(), eig<T>(), B.e_d.p, c<T>(), fac<T>(), in, base, out,
                                  sign, 1, act, s, sizeof(T) == 8 ? B.tuning.diag64_cpw : 0);
        else
            launch_dinv<T>(lay<T>(), q<T>(), qt<T>(), t.lam.p, B.e_d.p, c<T>(), fac<T>(), in, base,
                           out, sign, act, s);
    }
    // raw / raw_norm (DFT form only): in = raw / raw_norm[b] is formed (and stored to in) by
    // the first coupling; raw may alias out
    template <typename T>
    // shifted: out = F in with A = I + shifted_scale() F (F = K, -/+K^(order+1)): no identity
    // term, so the Arnoldi on F keeps the bits that A = I + small would cancel, and no D^-1
    // launch reads a base vector
    void jac(const cx<T> *in, cx<T> *out, const int *act, cx<T> *u, cx<T> *w1buf,
             const cx<T> *raw = nullptr, const double *raw_norm = nullptr, bool shifted = false) {
        auto first_coupling = [&](const cx<T> *src, cx<T> *dst) {
            if (raw) couple_full(raw, dst, act, raw_norm, const_cast<cx<T> *>(in));
            else couple_full(src, dst, act);
        };
        if (!neumann || order <= 1) {
            cx<T> *w1 = neumann ? w1buf : out;
            first_coupling(in, u);
            dinv<T>(u, shifted ? nullptr : in, w1, T(1), act);
            if (neumann) {
                couple_full(w1, u, act);
                if (shifted) dinv<T>(u, nullptr, out, T(1), act);
                else dinv<T>(u, w1, out, T(-1), act);
            }
            return;
        }
        // M (I + K) = I - (-K)^(order+1), K = D^-1 V: order + 1 products of K, then one sum
        const cx<T> *t = in;
        for (int i = 0; i < order; ++i) {
            if (i == 0) first_coupling(t, u);
            else couple_full(t, u, act);
            dinv<T>(u, nullptr, w1buf, T(1), act);
            t = w1buf;
        }
        couple_full(t, u, act);
        if (shifted) dinv<T>(u, nullptr, out, T(1), act);
        else dinv<T>(u, in, out, (order & 1) ? T(-1) : T(1), act);
    }
    double shifted_scale() const { return neumann && (order & 1) ? -1.0 : 1.0; }
    // s = M r = sum_{i<=order} (-K)^i r by Horner (s <- r - K s); `first` is the first
    // coupling product V r (null: computed here). s must not alias r.
    template <typename T>
    void neumann_rhs(const cx<T> *r, cx<T> *s_out, const int *act, cx<T> *u, const cx<T> *vr) {
        if (!neumann) {  // no factor: the right-hand side is r itself
            MSX_HIP(hipMemcpyAsync(s_out, r, sizeof(cx<T>) * B.n_elem(), hipMemcpyDeviceToDevice, s));
            return;
        }
        const cx<T> *cur = r;
        for (int i = 0; i < order; ++i) {
            if (i == 0 && vr) {
                dinv<T>(vr, r, s_out, T(-1), act);
            } else {
                couple_full(cur, u, act);
                dinv<T>(u, r, s_out, T(-1), act);
            }
            cur = s_out;
        }
    }
    // D^-1 src in FP64 for every instance; for the scattering source (nonzero only at node N
    // of the specular channels) by Sherman-Morrison on the specular channels alone
    void dinv_source(const cx<double> *src, cx<double> *out, bool scatter_source) {
        if (scatter_source) {
            ProfScope ps("specular_dinv_f64", s, 0.0, (double)B.n_elem() * sizeof(cx<double>));
            launch_specular_dinv(L, B.spec_d.p, B.bsrc_d.p, B.c_d.p, B.fac_d.p, out, s);
        } else {
            dinv<double>(src, nullptr, out, 1.0, nullptr);
        }
    }
    // Block Jacobi right-hand side in FP64: rhs = M D^-1 src
    void jac_rhs(const cx<double> *src, cx<double> *rhs, bool scatter_source) {
        cx<double> *r1 = neumann ? B.tmp2.p : rhs;
        dinv_source(src, r1, scatter_source);
        if (neumann) neumann_rhs<double>(r1, rhs, nullptr, B.tmp1.p, nullptr);
    }
    template <typename T> const T *q() const;
    template <typename T> const T *qt() const;
    template <typename T> EigTables<T> eig() const;
    template <typename T> const cx<T> *c() const;
    template <typename T> const T *fac() const;
};
template <> const double *Ops::q<double>() const { return t.Q.p; }
template <> const float *Ops::q<float>() const { return t.Qf.p; }
template <> const double *Ops::qt<double>() const { return t.QT.p; }
template <> const float *Ops::qt<float>() const { return t.QTf.p; }
template <> EigTables<double> Ops::eig<double>() const { return t.eig64(); }
template <> EigTables<float> Ops::eig<float>() const { return t.eig32(); }
template <> const cx<double> *Ops::c<double>() const { return B.c_d.p; }
template <> const cx<float> *Ops::c<float>() const { return B.cf_d.p; }
template <> const double *Ops::fac<double>() const { return B.fac_d.p; }
template <> const float *Ops::fac<float>() const { return L32.nm ? B.facf_nm_d.p : B.facf_d.p; }

// Krylov dimension of a solve. restart > 0: that dimension. restart = 0 (unrestarted, the
// reference's setting): maxiter, as long as the basis (n_elem complex values of elem_bytes
// per vector, allocated in blocks as the iterations need them) and the per-instance
// Hessenberg matrices ((k+1)^2 complex FP64 each, allocated up front) fit in a quarter of the
// device memory free now; past that the largest dimension that fits, after which GMRES
// restarts (at the benchmark's 256 conditions, 110,498 channels and nz = 150 this is ~260).
int krylov_dim(const msx_solve_options &o, int n_inst, long long n_elem, size_t elem_bytes) {
    int m = o.restart > 0 ? o.restart : o.maxiter;
    m = std::max(1, std::min(m, KR_MAX_BASIS - 1));
    if (o.restart <= 0) {
        size_t free_b = 0, total_b = 0;
        if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) {
            (void)hipGetLastError();
            free_b = (size_t)64 << 30;
        }
        const double budget = 0.25 * (double)free_b;
        auto need = [&](double k) {
            return (k + 1.0) * (double)n_elem * (double)elem_bytes +
                   (double)n_inst * (k + 1.0) * (k + 1.0) * 16.0;
        };
        while (m > 16 && need(m) > budget) m = std::max(16, (int)(m * 0.9));
    }
    return m;
}

void check_options(const msx_solve_options *o) {
    if (!o) throw Error(MSX_ERR_INVALID, "null options");
    if (o->method != MSX_METHOD_GAUSS_SEIDEL && o->method != MSX_METHOD_JACOBI)
        throw Error(MSX_ERR_INVALID, "unknown method");
    if (o->precision != MSX_PRECISION_FP64 && o->precision != MSX_PRECISION_MIXED)
        throw Error(MSX_ERR_INVALID, "unknown precision");
    if (o->precision == MSX_PRECISION_MIXED && o->method != MSX_METHOD_JACOBI)
        throw Error(MSX_ERR_INVALID, "mixed precision runs with the block-Jacobi method");
    if (!(o->eps > 0)) throw Error(MSX_ERR_INVALID, "eps must be positive");
    if (o->coupling != 0 && o->coupling != 1) throw Error(MSX_ERR_INVALID, "unknown coupling");
    if (o->diag < 0 || o->diag > 2) throw Error(MSX_ERR_INVALID, "unknown diag kernel");
    if (o->maxiter <= 0) throw Error(MSX_ERR_INVALID, "maxiter must be positive");
    if (o->window <= 0) throw Error(MSX_ERR_INVALID, "window must be positive");
    if (o->warm_start != 0 && o->warm_start != 1) throw Error(MSX_ERR_INVALID, "warm_start is 0 or 1");
    if (o->neumann < 0) throw Error(MSX_ERR_INVALID, "neumann must be a non-negative degree");
    if (o->neumann > 1 && o->method != MSX_METHOD_JACOBI)
        throw Error(MSX_ERR_INVALID, "a Neumann factor of degree > 1 needs the block-Jacobi method");
    if (o->inner_orth < 0 || o->inner_orth > 2) throw Error(MSX_ERR_INVALID, "inner_orth is 0, 1 or 2");
    if (o->autotune != 0 && o->autotune != 1) throw Error(MSX_ERR_INVALID, "autotune is 0 or 1");
    if (o->graphs != 0 && o->graphs != 1) throw Error(MSX_ERR_INVALID, "graphs is 0 or 1");
    if (o->inner_products < 0 || o->inner_products > 2)
        throw Error(MSX_ERR_INVALID, "inner_products is 0, 1 or 2");
    if (!(o->inner_eps >= 0 && o->inner_eps < 1))
        throw Error(MSX_ERR_INVALID, "inner_eps is 0 (automatic) or in (0, 1)");
}

// ---- autotuner ----------------------------------------------------------------------------
// At a batch's first solve with autotune set, times the launch-shape choices of its main
// kernels on its own sizes and keeps the fastest: the DFT coupling's nodes per wavefront (FP32
// for the mixed solver, FP64), the Arnoldi vector kernels' block length (the precision the
// Krylov basis runs in) and the FP64 diagonal blocks' channels per wavefront. The node and
// channel groupings do not change the arithmetic; the block length changes only the order in
// which the dot products' partial sums add up. Every shape is timed as a solve launches it with
// all instances iterating (the coupling's fall-back to one node per wavefront when a grouping
// leaves the device short of wavefronts included). Choices are cached per batch shape (table,
// nz, instances, channel counts, vector order) for the process, and in the text file named by
// MSX_TUNE_FILE when that is set (one line per shape: key and the five values).
std::mutex g_tune_mu;
std::map<std::string, msx_batch::Tuning> g_tune_cache;
bool g_tune_file_read = false;

std::string tune_key(const msx_batch &B, const Ops &ops, const msx_solve_options &o) {
    char buf[320];
    std::snprintf(buf, sizeof buf,
                  "nkx=%d,nky=%d,nz=%d,inst=%d,channels=%lld,max_nc=%d,nm=%d%d,direct=%d,vdiag=%d,prec=%d",
                  B.t->nkx, B.t->nky, B.t->nz, B.n_inst, B.total_ch, B.max_nc, ops.L32.nm,
                  ops.L64.nm, ops.direct ? 1 : 0, ops.vector_diag ? 1 : 0, o.precision);
    return buf;
}

void read_tune_file() {
    if (g_tune_file_read) return;
    g_tune_file_read = true;
    const char *path = std::getenv("MSX_TUNE_FILE");
    if (!path || !*path) return;
    std::ifstream in(path);
    std::string key;
    msx_batch::Tuning t;
    // one shape a line: key and the choices; a line without the last field (written before
    // it existed) leaves it at its built-in value
    std::string line;
    while (std::getline(in, line)) {
        std::istringstream ls(line);
        msx_batch::Tuning t;
        if (!(ls >> key >> t.dft_npw32 >> t.dft_npw64 >> t.chunk32 >> t.chunk64 >> t.diag64_cpw))
            continue;
        if (!(ls >> t.diag32_cpw)) t.diag32_cpw = 0;
        t.done = true;
        t.key = key;
        g_tune_cache[key] = t;
    }
}

void append_tune_file(const msx_batch::Tuning &t) {
    const char *path = std::getenv("MSX_TUNE_FILE");
    if (!path || !*path) return;
    std::ofstream out(path, std::ios::app);
    out << t.key << ' ' << t.dft_npw32 << ' ' << t.dft_npw64 << ' ' << t.chunk32 << ' '
        << t.chunk64 << ' ' << t.diag64_cpw << ' ' << t.diag32_cpw << '\n';
}

// Index of the shape to keep among measured times ts: the built-in one (index def) unless
// another is faster by more than 3%. Launch-time differences below that are within what the
// same kernel varies between runs, and the built-in shapes were chosen on whole solves, where
// the active set shrinks as conditions converge (a shape timed with every instance active can
// lose there: eight nodes per wavefront measured 0.2% faster alone and 8% slower in a solve).
int pick_with_margin(const std::vector<double> &ts, int def) {
    int best = def;
    for (int i = 0; i < (int)ts.size(); ++i)
        if (ts[i] < ts[best] && ts[i] < 0.97 * ts[def]) best = i;
    return best;
}

// Shortest of `reps` timed runs of f (after one untimed run), in microseconds.
struct LaunchTimer {
    hipStream_t s;
    hipEvent_t a = nullptr, b = nullptr;
    explicit LaunchTimer(hipStream_t st) : s(st) {
        MSX_HIP(hipEventCreate(&a));
        MSX_HIP(hipEventCreate(&b));
    }
    ~LaunchTimer() {
        if (a) (void)hipEventDestroy(a);
        if (b) (void)hipEventDestroy(b);
    }
    template <typename F>
    double best_us(F f, int reps = 3) {
        f();
        double best = 1e300;
        for (int r = 0; r < reps; ++r) {
            MSX_HIP(hipEventRecord(a, s));
            f();
            MSX_HIP(hipEventRecord(b, s));
            MSX_HIP(hipEventSynchronize(b));
            float ms = 0;
            MSX_HIP(hipEventElapsedTime(&ms, a, b));
            best = std::min(best, 1e3 * (double)ms);
        }
        return best;
    }
};

// Arnoldi block length: the two main passes of an iteration (6 basis vectors: dot products
// with their reduction, then the update with the norm partials) at each candidate.
template <typename T>
long long tune_chunk(msx_batch &B, LaunchTimer &tm, std::string &rep) {
    const int kc = 6;
    const size_t n = (size_t)B.n_elem();
    const std::vector<long long> start = B.kstart(), len = B.klen();
    long long maxlen = 1;
    for (long long l : len) maxlen = std::max(maxlen, l);
    DBuf<long long> start_d(B.n_inst), len_d(B.n_inst);
    start_d.upload(start.data(), B.n_inst, tm.s);
    len_d.upload(len.data(), B.n_inst, tm.s);
    std::vector<int> ones(B.n_inst, 1);
    DBuf<int> act(B.n_inst);
    act.upload(ones.data(), B.n_inst, tm.s);
    DBuf<cx<T>> vecs((size_t)(kc + 1) * n);
    MSX_HIP(hipMemsetAsync(vecs.p, 0, sizeof(cx<T>) * (size_t)(kc + 1) * n, tm.s));
    std::vector<cx<T> *> vp(kc);
    for (int j = 0; j < kc; ++j) vp[j] = vecs.p + (size_t)j * n;
    DBuf<cx<T> *> vp_d(kc);
    vp_d.upload(vp.data(), kc, tm.s);
    cx<T> *w = vecs.p + (size_t)kc * n;
    const long long cands[] = {1024, 2048, 4096, 8192};
    const size_t slots = (size_t)B.n_inst * (kc + 1);
    DBuf<double> partial(slots * (size_t)((maxlen + 1023) / 1024) * 2);
    DBuf<cx<double>> h(slots);
    std::vector<double> ts;
    rep += "\"arnoldi_block_us_" + std::string(sizeof(T) == 4 ? "f32" : "f64") + "\": {";
    for (size_t i = 0; i < sizeof(cands) / sizeof(cands[0]); ++i) {
        KLayout K{};
        K.n_inst = B.n_inst;
        K.kmax = kc;
        K.chunk = cands[i];
        K.nchunk = (int)((maxlen + K.chunk - 1) / K.chunk);
        K.start = start_d.p;
        K.len = len_d.p;
        K.active = act.p;
        const double t = tm.best_us([&] {
            k_dots<T>(K, vp_d.p, kc, w, 0, partial.p, h.p, 0, tm.s);
            k_axpy<T>(K, vp_d.p, kc, h.p, w, 1, partial.p, h.p, tm.s);
        });
        rep += (i ? ", \"" : "\"") + std::to_string(cands[i]) + "\": " + std::to_string(t);
        ts.push_back(t);
    }
    rep += "}, ";
    MSX_HIP(hipStreamSynchronize(tm.s));
    return cands[pick_with_margin(ts, 1)];
}

// The batch's Arnoldi block lengths: the first value tuned for each precision is kept and
// replaces what a later key of the same batch brings (the block length depends on the batch's
// vector lengths only, not on the vector order or the operator form); the precision a key does
// not tune takes the kept value too, if there is one.
void keep_chunks(msx_batch &B, bool mixed) {
    long long &kept = mixed ? B.kept_chunk32 : B.kept_chunk64;
    long long &cur = mixed ? B.tuning.chunk32 : B.tuning.chunk64;
    if (kept) cur = kept;
    else kept = cur;
    const long long other = mixed ? B.kept_chunk64 : B.kept_chunk32;
    if (other) (mixed ? B.tuning.chunk64 : B.tuning.chunk32) = other;
}

void autotune(msx_batch &B, Ops &ops, const msx_solve_options &o) {
    const std::string key = tune_key(B, ops, o);
    {
        std::lock_guard<std::mutex> lock(g_tune_mu);
        read_tune_file();
        auto it = g_tune_cache.find(key);
        if (it != g_tune_cache.end()) {
            B.tuning = it->second;
            keep_chunks(B, o.precision == MSX_PRECISION_MIXED);
            B.tuning.report = "{\"source\": \"cache\", \"key\": \"" + key + "\", \"dft_nodes_per_wave_f32\": " +
                              std::to_string(B.tuning.dft_npw32) + ", \"dft_nodes_per_wave_f64\": " +
                              std::to_string(B.tuning.dft_npw64) + ", \"arnoldi_block_f32\": " +
                              std::to_string(B.tuning.chunk32) + ", \"arnoldi_block_f64\": " +
                              std::to_string(B.tuning.chunk64) + ", \"diag_f64_channels_per_wave\": " +
                              std::to_string(B.tuning.diag64_cpw) + ", \"diag_split_channels_per_wave\": " +
                              std::to_string(B.tuning.diag32_cpw ? B.tuning.diag32_cpw : 16) + "}";
            return;
        }
    }
    hipStream_t s = B.t->stream;
    const size_t n = (size_t)B.n_elem();
    LaunchTimer tm(s);
    msx_batch::Tuning T;
    T.key = key;
    std::string rep = "{\"source\": \"measured\", \"key\": \"" + key + "\", ";
    const bool mixed = o.precision == MSX_PRECISION_MIXED;
    const int npw_c[] = {1, 2, 4, 8};
    if (!ops.direct) {
        msx_batch::need(B.tmp1, n);
        msx_batch::need(B.tmp2, n);
        std::vector<double> ts;
        rep += "\"dft_us_f64\": {";
        for (int i = 0; i < 4; ++i) {
            ts.push_back(tm.best_us([&] {
                launch_dft_coupling<double>(ops.L64, ops.dft->d64(), B.tmp1.p, B.tmp2.p, nullptr, s,
                                            B.n_inst, nullptr, nullptr, npw_c[i]);
            }));
            rep += (i ? ", \"" : "\"") + std::to_string(npw_c[i]) + "\": " + std::to_string(ts[i]);
        }
        T.dft_npw64 = npw_c[pick_with_margin(ts, 2)];
        rep += "}, ";
        if (mixed) {
            msx_batch::need(B.f_tmp1, n);
            msx_batch::need(B.f_tmp2, n);
            std::vector<double> ts;
            rep += "\"dft_us_f32\": {";
            for (int i = 0; i < 4; ++i) {
                ts.push_back(tm.best_us([&] {
                    launch_dft_coupling<float>(ops.L32, ops.dft->d32(), B.f_tmp1.p, B.f_tmp2.p,
                                               nullptr, s, B.n_inst, nullptr, nullptr, npw_c[i]);
                }));
                rep += (i ? ", \"" : "\"") + std::to_string(npw_c[i]) + "\": " + std::to_string(ts[i]);
            }
            T.dft_npw32 = npw_c[pick_with_margin(ts, 2)];
            rep += "}, ";
        }
    }
    if (!ops.vector_diag && diag_mfma_supported(B.t->nz) && B.t->nzp <= 160) {
        msx_batch::need(B.tmp1, n);
        msx_batch::need(B.tmp2, n);
        const int cpw_c[] = {8, 16};
        std::vector<double> ts;
        rep += "\"diag_us_f64\": {";
        for (int i = 0; i < 2; ++i) {
            ts.push_back(tm.best_us([&] {
                launch_diag_blocks<double>(ops.L, B.t->eig64(), B.e_d.p, B.c_d.p, B.fac_d.p, B.tmp1.p,
                                           nullptr, B.tmp2.p, 1.0, 1, nullptr, s, cpw_c[i]);
            }));
            rep += (i ? ", \"" : "\"") + std::to_string(cpw_c[i]) + "\": " + std::to_string(ts[i]);
        }
        T.diag64_cpw = cpw_c[pick_with_margin(ts, 1)];
        rep += "}, ";
    }
    // the split-FP16 FP32 D^-1 (mixed precision, node-major vectors): 16 or 8 channels per
    // wavefront (two or four wavefronts per SIMD)
    if (mixed && ops.split_f16 && !ops.vector_diag && ops.L32.nm &&
        diag_h3_fits(ops.L32, B.t->h3())) {
        msx_batch::need(B.f_tmp1, n);
        msx_batch::need(B.f_tmp2, n);
        const int cpw_c[] = {16, 8};
        std::vector<double> ts;
        rep += "\"diag_split_us_f32\": {";
        for (int i = 0; i < 2; ++i) {
            ts.push_back(tm.best_us([&] {
                launch_diag_h3(ops.L32, B.t->h3(), B.t->lam.p, B.e_d.p, B.cf_d.p, B.facf_nm_d.p,
                               B.f_tmp1.p, nullptr, B.f_tmp2.p, 1.0f, nullptr, s, cpw_c[i]);
            }));
            rep += (i ? ", \"" : "\"") + std::to_string(cpw_c[i]) + "\": " + std::to_string(ts[i]);
        }
        T.diag32_cpw = cpw_c[pick_with_margin(ts, 0)];
        rep += "}, ";
    }
    if (mixed) T.chunk32 = B.kept_chunk32 ? B.kept_chunk32 : tune_chunk<float>(B, tm, rep);
    else T.chunk64 = B.kept_chunk64 ? B.kept_chunk64 : tune_chunk<double>(B, tm, rep);
    rep += "\"dft_nodes_per_wave_f32\": " + std::to_string(T.dft_npw32) +
           ", \"dft_nodes_per_wave_f64\": " + std::to_string(T.dft_npw64) +
           ", \"arnoldi_block_f32\": " + std::to_string(T.chunk32) +
           ", \"arnoldi_block_f64\": " + std::to_string(T.chunk64) +
           ", \"diag_f64_channels_per_wave\": " + std::to_string(T.diag64_cpw) +
           ", \"diag_split_channels_per_wave\": " + std::to_string(T.diag32_cpw ? T.diag32_cpw : 16) + "}";
    T.done = true;
    T.report = rep;
    B.tuning = T;
    keep_chunks(B, mixed);
    std::lock_guard<std::mutex> lock(g_tune_mu);
    g_tune_cache[key] = T;
    append_tune_file(T);
}

// Solve A psi = src for all instances into `x` (FP64). transposed: A^T with the flipped table.
// scatter_source: the source is the scattering source (b at node N of the specular channels);
// the block-Jacobi paths then form D^-1 src from the channel tables alone, and src_in (the
// vector launch_set_source writes) is read only by Gauss-Seidel and may be null otherwise.
// src_in is only read.
void solve_system(msx_batch &B, const msx_solve_options &o, const cx<double> *src_in,
                  cx<double> *x, bool transposed, std::vector<int> &iters,
                  std::vector<int> &info, std::vector<double> &resid, bool scatter_source) {
    msx_table &t = *B.t;
    hipStream_t s = t.stream;
    const size_t n = (size_t)B.n_elem();
    msx_batch::need(B.src, n);
    msx_batch::need(B.rhs, n);
    msx_batch::need(B.tmp1, n);
    msx_batch::need(B.tmp2, n);
    if (!B.factors_ready) B.build_factors();
    Ops ops(B, o.neumann >= 1, transposed, o.coupling == 1);
    ops.order = std::max(1, o.neumann);
    ops.vector_diag = o.diag == 1;  // auto: matrix cores when nz <= 320
    ops.split_f16 = o.inner_products != 1;
    ops.split_f16_required = o.inner_products == 2;
    ops.set_fp32_order();
    // The mixed solver's refinement residual as D^-1 (s - D x - V x) with D^-1 in FP32
    // (split_residual) instead of d0 - x - D^-1 V x in FP64: for the scattering source or a
    // given one (the adjoint's: D is complex symmetric, and the transposed coupling is the
    // flipped table's) at tolerances down to 1e-9, where the FP64 cancellation in s - D x - V x
    // (about 1e-16 |D| |x|) stays far below the residuals the test has to resolve
    const bool mixed = o.precision == MSX_PRECISION_MIXED;
    const bool split = mixed && (scatter_source || src_in) && o.eps >= 1e-9 &&
                       split_residual_supported(t.nzp) && t.nzp >= t.nz &&
                       (scatter_source || o.method == MSX_METHOD_JACOBI);
    // With the split residual, the scattering source and node-major FP32 vectors, the FP64
    // refinement vectors (the state, D^-1 s, the coupling of the state) are node-major too: the
    // FP64 coupling, the split residual and the solution update then read and write whole
    // lines. The batch's state is left in that order (psi_nm). A given source stays in the
    // batch order, and so does everything of its solve.
    const bool nm64 = split && scatter_source && !transposed && ops.L32.nm && x == B.psi_d.p;
    if (nm64) ops.L64.nm = 1;
    if (o.autotune == 1 && o.method == MSX_METHOD_JACOBI &&
        tune_key(B, ops, o) != B.tuning.key)
        autotune(B, ops, o);
    const cx<double> *src = src_in;
    if (!src && !(scatter_source && o.method == MSX_METHOD_JACOBI))
        throw Error(MSX_ERR_INVALID, "solve: no source vector");
    // A warm start (forward solves only) keeps x as the initial guess, in the order this path
    // runs in.
    const bool warm = o.warm_start == 1 && !transposed;
    if (x == B.psi_d.p) {
        if (warm) B.psi_order(nm64);
        else B.psi_nm = nm64;
    }
    if (!warm) MSX_HIP(hipMemsetAsync(x, 0, sizeof(cx<double>) * n, s));
    // sized once per batch and option pair (free memory shrinks once the basis exists, and a
    // dimension that moved with it would reallocate the workspace on every solve)
    auto &memo = B.kdim_memo[o.precision == MSX_PRECISION_FP64 ? 0 : 1];
    if (memo.restart != o.restart || memo.maxiter != o.maxiter) {
        memo.dim = krylov_dim(o, B.n_inst, B.n_elem(),
                              o.precision == MSX_PRECISION_FP64 ? sizeof(cx<double>)
                                                                : sizeof(cx<float>));
        memo.restart = o.restart;
        memo.maxiter = o.maxiter;
    }
    const int kdim = memo.dim;
    iters.assign(B.n_inst, 0);
    info.assign(B.n_inst, 0);
    resid.assign(B.n_inst, 0.0);

    if (o.precision == MSX_PRECISION_FP64) {
        if (!B.kr64 || B.kr64->kmax != std::min(kdim, KR_MAX_BASIS - 1) ||
            B.kr64->K.chunk != B.tuning.chunk64)
            B.kr64.reset(new Krylov<double>(B.kstart(), B.klen(), kdim, s, B.tuning.chunk64));
        Krylov<double> &kr = *B.kr64;
        kr.set_all_active(nullptr);
        Op<double> op;
        if (o.method == MSX_METHOD_GAUSS_SEIDEL) {
            // the right-hand side is formed in place (L^-1 src): on a copy, src_in is only read
            MSX_HIP(hipMemcpyAsync(B.src.p, src, sizeof(cx<double>) * n, hipMemcpyDeviceToDevice, s));
            ops.gs_rhs(B.src.p, B.rhs.p);
            op = [&](const cx<double> *in, cx<double> *out, const int *act) { ops.gs(in, out, act); };
        } else {
            ops.jac_rhs(src, B.rhs.p, scatter_source);
            op = [&](const cx<double> *in, cx<double> *out, const int *act) {
                ops.jac<double>(in, out, act, B.tmp1.p, B.tmp2.p);
            };
        }
        ops.snap_src = &kr.snap; ops.count_src = &kr.hint_count;
        // block Jacobi with the DFT coupling: new basis vectors normalised by the operator
        const OpScaled<double> op_s = [&](const cx<double> *raw, const double *nrm, cx<double> *in,
                                          cx<double> *out, const int *act) {
            ops.jac<double>(in, out, act, B.tmp1.p, B.tmp2.p, raw, nrm);
        };
        const bool fused = o.method == MSX_METHOD_JACOBI && !ops.direct;
        const std::string sig64 = ops.graph_signature(
            o.method == MSX_METHOD_JACOBI ? "jacobi64" : "gs64", B.tmp1.p, B.tmp2.p, B.rhs.p);
        const std::function<int()> shape64 = [&] { return ops.launch_shape_key(); };
        // block Jacobi: the selective second Gram-Schmidt pass; Gauss-Seidel keeps two passes
        // (its iteration counts are the reference's)
        gmres<double, double>(kr, op, B.rhs.p, x, !warm, o.eps, o.window, o.maxiter, true,
                              o.method == MSX_METHOD_JACOBI ? 3 : 2,
                              nullptr, nullptr, fused ? &op_s : nullptr,
                              o.graphs == 1 ? &sig64 : nullptr, &shape64);
        ops.snap_src = nullptr; ops.count_src = nullptr;
        iters = kr.host_ints(kr.iters);
        auto conv = kr.host_ints(kr.converged);
        resid = kr.host_doubles(kr.resid);
        for (int b = 0; b < B.n_inst; ++b) info[b] = conv[b] ? 0 : 1;
        return;
    }

    // Mixed precision (block Jacobi): iterative refinement of (I + D^-1 V) x = D^-1 s. Each
    // sweep forms the residual r = D^-1 s - x - D^-1 V x in FP64 (one coupling and one D^-1),
    // stops an instance when |r| < eps |D^-1 s|, and otherwise solves the correction equation
    // M (I + D^-1 V) d = M r with FP32 GMRES on the matrix cores (M = I - D^-1 V, the Neumann
    // factor, applied to r in FP32); the corrections accumulate in FP64.
    msx_batch::need(B.tmp3, n);
    msx_batch::need(B.f_rhs, n);
    msx_batch::need(B.f_tmp1, n);
    msx_batch::need(B.f_tmp2, n);
    cx<double> *d0 = B.rhs.p;  // D^-1 s
    if (nm64) {
        // node-major, with its FP32 copy (the inner right-hand side of a cold first sweep)
        ProfScope ps("specular_dinv_f64", s, 0.0, (double)B.n_elem() * (sizeof(cx<double>) + sizeof(cx<float>)));
        launch_specular_dinv(ops.L64, B.spec_d.p, B.bsrc_d.p, B.c_d.p, B.fac_d.p, d0, s, B.f_tmp2.p);
    } else {
        ops.dinv_source(src, d0, scatter_source);
    }
    if (!B.kr_outer || B.kr_outer->K.chunk != B.tuning.chunk64)
        B.kr_outer.reset(new Krylov<double>(B.kstart(), B.klen(), 1, s, B.tuning.chunk64));
    Krylov<double> &outer = *B.kr_outer;
    outer.set_all_active(nullptr);
    const int inner_dim = std::max(2, std::min(kdim, 200));
    if (!B.kr32 || B.kr32->kmax != inner_dim || B.kr32->K.chunk != B.tuning.chunk32)
        B.kr32.reset(new Krylov<float>(B.kstart(), B.klen(), inner_dim, s, B.tuning.chunk32));
    Krylov<float> &inner = *B.kr32;
    // the inner Arnoldi runs on F = (A - I) / scale (the operator without its identity term)
    inner.S.hscale = ops.shifted_scale();
    inner.S.hshift = 1.0;
    Op<float> op32 = [&](const cx<float> *in, cx<float> *out, const int *act) {
        ops.jac<float>(in, out, act, B.f_tmp1.p, B.f_tmp2.p, nullptr, nullptr, true);
    };
    const OpScaled<float> op32s = [&](const cx<float> *raw, const double *nrm, cx<float> *in,
                                      cx<float> *out, const int *act) {
        ops.jac<float>(in, out, act, B.f_tmp1.p, B.f_tmp2.p, raw, nrm, true);
    };
    // |D^-1 s|^2 per instance: on a cold start it is also the first residual's
    const std::vector<double> rhs_n2 = outer.norms2(d0);
    std::vector<int> act(B.n_inst, 1);
    msx_batch::need(B.iters_acc, (size_t)B.n_inst);
    MSX_HIP(hipMemsetAsync(B.iters_acc.p, 0, sizeof(int) * B.n_inst, s));
    // unset: loose enough that two refinement sweeps (the residual shrinking by about
    // inner_eps each) meet eps with a factor ~10 to spare, never tighter than FP32 GMRES
    // reliably reaches (1e-5) nor looser than 2.5e-4 (at the benchmark's eps 5e-7: 2.4e-4,
    // |dP| 2.6e-8 against a tight FP64 solve, 3% faster than 1e-4)
    const double inner_eps =
        o.inner_eps > 0 ? o.inner_eps : std::min(2.5e-4, std::max(1e-5, std::sqrt(o.eps) / 3.0));
    const int max_refine = o.max_refine > 0 ? o.max_refine : 6;
    cx<double> *r64 = B.tmp3.p;
    // per-instance inner tolerances (host copy; uploaded before each inner solve)
    std::vector<double> eps_in(B.n_inst, inner_eps);
    msx_batch::need(B.eps_in_d, (size_t)B.n_inst);
    constexpr double kSweepMargin = 0.03;
    // Gram-Schmidt passes of the inner Arnoldi: by default one classical pass, and the second
    // only for the instances whose first pass left the new vector's projection on the basis
    // above 1e-4 of its norm (measured from the second projection, which the first update's
    // pass over the basis forms anyway), for Neumann degrees up to 2; two passes above (the
    // high powers of D^-1 V spread the operator's spectrum: at degree 7 the selective form left
    // a condition unconverged); 1: one pass; 2: two
    const int orth = o.inner_orth == 1   ? 1
                     : o.inner_orth == 2 ? 2
                     : (!ops.neumann || ops.order <= 2) ? 3 : 2;
    for (int sweep = 0;; ++sweep) {
        outer.set_all_active(act.data());
        // a cold start's first residual is d0 itself
        const cx<double> *res64 = (sweep == 0 && !warm) ? d0 : r64;
        bool r_in_fp32 = false;  // the residual is already in B.f_tmp2 (FP32, inner order)
        std::vector<double> r_n2;
        if (sweep == 0 && !warm) {
            r_n2 = rhs_n2;
            r_in_fp32 = nm64;
        } else if (split) {
            // r = D^-1 rho, rho = s - D x - V x in FP64 (one FP64 coupling, one FP64 product
            // with H0), D^-1 rho in FP32: rho is small, so FP32 holds r to ~1e-6 relative
            ops.snap_src = &outer.snap; ops.count_src = &outer.hint_count;
            ops.couple_full(x, B.tmp1.p, outer.active.p);
            {
                Work fl, by;
                if (profiling()) {
                    const double nzd = t.nz;
                    fl = ops.over_active(outer.active.p, [&](int nc) { return nc * (4.0 * nzd * nzd + 16.0 * nzd); });
                    const double per = scatter_source ? 40.0 : 56.0;  // + the source vector
                    by = ops.over_active(outer.active.p, [&](int nc) { return nc * nzd * per; });
                }
                ProfScope ps("residual_split_f64", s, fl, by);
                launch_split_residual(ops.L64, ops.L32, t.H0p.p, t.nzp, B.e_d.p, B.c_d.p, B.spec_d.p,
                                      B.bsrc_d.p, x, B.tmp1.p, B.f_tmp1.p, outer.active.p, s,
                                      scatter_source ? nullptr : src);
            }
            ops.dinv<float>(B.f_tmp1.p, nullptr, B.f_tmp2.p, 1.0f, outer.active.p);
            inner.set_all_active(act.data());
            r_n2 = inner.norms2(B.f_tmp2.p);
            r_in_fp32 = true;
        } else {
            // r = d0 - (x + D^-1 V x)   (B.src serves as scratch)
            ops.snap_src = &outer.snap; ops.count_src = &outer.hint_count;
            ops.couple_full(x, B.tmp1.p, outer.active.p);
            ops.dinv<double>(B.tmp1.p, x, B.src.p, 1.0, outer.active.p);
            k_axpby<double>(outer.K, d0, B.src.p, -1.0, r64, outer.active.p, s);
            r_n2 = outer.norms2(r64);
        }
        int n_act = 0;
        for (int b = 0; b < B.n_inst; ++b) {
            if (!act[b]) continue;
            const double rel = rhs_n2[b] > 0 ? std::sqrt(r_n2[b] / rhs_n2[b]) : 0.0;
            resid[b] = rel;
            // the inner solve need only shrink this residual to a small fraction of eps
            // (relative to |D^-1 s|): never by more than inner_eps, nor by less than a factor
            // 10. The margin leaves room for the amplification between the inner (Neumann-
            // preconditioned) and the outer residual: at 1/3 and 1/10 some conditions took a
            // third sweep, whose nearly empty GMRES costs more than the iterations saved
            eps_in[b] = rel > 0 ? std::min(0.1, std::max(inner_eps, kSweepMargin * o.eps / rel))
                                : inner_eps;
            if (rel < o.eps) act[b] = 0;
            else ++n_act;
        }
        if (n_act == 0 || sweep >= max_refine) break;
        B.eps_in_d.upload(eps_in.data(), (size_t)B.n_inst, s);
        outer.set_all_active(act.data());
        ops.snap_src = &outer.snap; ops.count_src = &outer.hint_count;
        // inner right-hand side M r in FP32 (M = I - D^-1 V to first degree; on a cold start the
        // first r = D^-1 s lives on the specular channels only, and V r is a single-source product)
        if (!r_in_fp32) {
            if (nm64) launch_to_float((long long)n, res64, B.f_tmp2.p, s);  // same order
            else launch_to_float_vec(ops.L32, (long long)n, res64, B.f_tmp2.p, s);
        }
        if (!ops.neumann) {
            // no Neumann factor: nothing to couple
        } else if (sweep == 0 && !warm && scatter_source && !transposed) {
            const double fl = 8.0 * (double)B.total_ch * t.nz;
            const double by = 2.0 * (double)B.total_ch * t.nz * sizeof(cx<float>);
            ProfScope ps("coupling_specular_f32", s, fl, by);
            launch_specular_coupling<float>(ops.L32, t.Vf.p, B.spec_d.p, B.f_tmp2.p, B.f_tmp1.p,
                                            outer.active.p, s);
        } else {
            ops.couple_full(B.f_tmp2.p, B.f_tmp1.p, outer.active.p);
        }
        ops.neumann_rhs<float>(B.f_tmp2.p, B.f_rhs.p, outer.active.p, B.f_tmp1.p,
                               ops.neumann ? B.f_tmp1.p : nullptr);
        inner.set_all_active(act.data());
        ops.snap_src = &inner.snap; ops.count_src = &inner.hint_count;
        const std::string sig32 = ops.graph_signature("mixed32", B.f_tmp1.p, B.f_tmp2.p, B.eps_in_d.p);
        const std::function<int()> shape32 = [&] { return ops.launch_shape_key(); };
        // The first sweep's inner solve ends for everyone once at most a tenth of a large batch
        // still iterates: its last iterations carry a few conditions at latency-bound cost, and
        // the next sweep, which every condition takes after a cold first sweep, continues them
        // from their current solution (21.6 -> 21.3 ms a step on the benchmark batch, |dP|
        // 2.6e-8 -> 8e-8 against the tight solve)
        const int stop_below = sweep == 0 && n_act >= 32 ? n_act / 10 : 0;
        gmres<float, double>(inner, op32, B.f_rhs.p, x, true, inner_eps, 1,
                             std::min(o.maxiter, inner_dim), false, orth,
                             nm64 ? nullptr : &ops.L32, B.eps_in_d.p, ops.direct ? nullptr : &op32s,
                             o.graphs == 1 ? &sig32 : nullptr, &shape32, stop_below);
        launch_add_ints(B.n_inst, inner.iters.p, B.iters_acc.p, s);
    }
    {
        std::vector<int> it(B.n_inst);
        B.iters_acc.download(it.data(), B.n_inst, s);
        MSX_HIP(hipStreamSynchronize(s));
        for (int b = 0; b < B.n_inst; ++b) iters[b] += it[b];
    }
    for (int b = 0; b < B.n_inst; ++b) info[b] = act[b] ? 1 : 0;
}

int fail(const Error &e) {
    set_last_error(e.what());
    return e.code;
}

// Column-major (nkx, nky, nz) complex -> row-major [g][z] with g = ix*nky + iy.
void table_from_fortran(int nkx, int nky, int nz, const double *src, std::vector<cx<double>> &dst) {
    dst.resize((size_t)nkx * nky * nz);
    for (int k = 0; k < nz; ++k)
        for (int iy = 0; iy < nky; ++iy)
            for (int ix = 0; ix < nkx; ++ix) {
                const size_t f = (size_t)ix + (size_t)nkx * ((size_t)iy + (size_t)nky * k);
                dst[((size_t)ix * nky + iy) * nz + k] = mk(src[2 * f], src[2 * f + 1]);
            }
}

void transpose_square(int n, const double *src, std::vector<double> &dst) {
    dst.resize((size_t)n * n);
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) dst[(size_t)i * n + j] = src[(size_t)j * n + i];
}

// Full-grid storage-order batch of one condition (the Fortran path's channel set).
msx_batch *full_grid_batch(msx_table *t, const double *e_col, const double *wave_b,
                           const double *wave_c, const double *wave_a) {
    const int nkx = t->nkx, nky = t->nky, nc = nkx * nky;
    std::vector<int> chx(nc), chy(nc);
    std::vector<double> e(nc), b(2 * (size_t)nc), c(2 * (size_t)nc);
    for (int g = 0; g < nc; ++g) {
        const int ix = g / nky, iy = g % nky;
        chx[g] = ix;
        chy[g] = iy;
        e[g] = e_col ? e_col[ix + (size_t)nkx * iy] : 0.0;
        b[2 * g] = wave_b ? wave_b[2 * g] : 0.0;
        b[2 * g + 1] = wave_b ? wave_b[2 * g + 1] : 0.0;
        c[2 * g] = wave_c ? wave_c[2 * g] : 0.0;
        c[2 * g + 1] = wave_c ? wave_c[2 * g + 1] : 0.0;
    }
    double a0[2] = {wave_a ? wave_a[0] : 0.0, wave_a ? wave_a[1] : 0.0};
    msx_batch *B = nullptr;
    int rc = msx_batch_create(t, 1, &nc, chx.data(), chy.data(), e.data(), b.data(), c.data(), a0,
                              &B);
    if (rc != MSX_OK) throw Error(rc, msx_last_error());
    return B;
}

struct TableGuard {
    msx_table *t = nullptr;
    ~TableGuard() { if (t) msx_table_destroy(t); }
};
struct BatchGuard {
    msx_batch *b = nullptr;
    ~BatchGuard() { if (b) msx_batch_destroy(b); }
};

}  // namespace

// =======================================================================================
// C ABI
// =======================================================================================
extern "C" {

const char *msx_last_error(void) { return g_last_error.c_str(); }

int msx_synchronize(void) {
    try {
        MSX_HIP(hipDeviceSynchronize());
        return MSX_OK;
    } catch (const Error &e) {
        return fail(e);
    }
}

int msx_device_info(char *buf, int buflen, int *device_count) {
    try {
        int n = 0;
        MSX_HIP(hipGetDeviceCount(&n));
        if (device_count) *device_count = n;
        if (n <= 0) throw Error(MSX_ERR_DEVICE, "no GPU visible");
        hipDeviceProp_t p;
        MSX_HIP(hipGetDeviceProperties(&p, 0));
        if (buf && buflen > 0) {
            std::strncpy(buf, p.gcnArchName, buflen - 1);
            buf[buflen - 1] = 0;
        }
        return MSX_OK;
    } catch (const Error &e) {
        return fail(e);
    }
}

void msx_default_options(msx_solve_options *opt, int nsig) {
    if (!opt) return;
    opt->method = MSX_METHOD_GAUSS_SEIDEL;
    opt->precision = MSX_PRECISION_FP64;
    opt->neumann = 1;
    opt->eps = 0.5 * std::pow(10.0, -(double)nsig);  // multiscat.f90:35
    opt->maxiter = 2000;                             // multiscat_gmres.f90:7,124
    opt->restart = 0;
    opt->window = 3;                                 // multiscat_gmres.f90:125
    opt->inner_eps = 0;  // sqrt(eps)/3 within [1e-5, 2.5e-4]
    opt->max_refine = 6;
    opt->coupling = 0;
    opt->diag = 0;
    opt->warm_start = 0;
    opt->inner_orth = 0;
    opt->autotune = 0;
    opt->graphs = 0;
    opt->inner_products = 0;
}

int msx_table_create(int nkx, int nky, int nz, const double *potential, const double *kinetic,
                     msx_table **out) {
    try {
        if (!out || !potential || !kinetic) throw Error(MSX_ERR_INVALID, "null argument");
        *out = nullptr;
        if (nkx <= 0 || nky <= 0 || nz <= 0) throw Error(MSX_ERR_DIMS, "non-positive dimension");
        if (nz < 2) throw Error(MSX_ERR_INVALID, "nz must be at least 2");
        std::unique_ptr<msx_table> t(new msx_table);
        t->nkx = nkx;
        t->nky = nky;
        t->nz = nz;
        MSX_HIP(hipStreamCreate(&t->stream));
        const size_t nG = (size_t)nkx * nky, n2 = (size_t)nz * nz;
        t->V.alloc(nG * nz);
        t->Vf.alloc(nG * nz);
        t->V.upload(reinterpret_cast<const cx<double> *>(potential), nG * nz, t->stream);
        t->T.alloc(n2);
        t->T.upload(kinetic, n2, t->stream);
        for (auto *b : {&t->A, &t->Vm, &t->Q, &t->QT}) b->alloc(n2);
        t->lam.alloc(nz);
        t->Qf.alloc(n2);
        t->QTf.alloc(n2);
        t->info.alloc(1);
        t->ticks.alloc(8);
        t->prepare();
        *out = t.release();
        return MSX_OK;
    } catch (const Error &e) {
        return fail(e);
    }
}

int msx_table_prepare(msx_table *t) {
    try {
        if (!t) throw Error(MSX_ERR_INVALID, "null table");
        t->prepare();
        return MSX_OK;
    } catch (const Error &e) {
        return fail(e);
    }
}

int msx_table_set_potential(msx_table *t, const double *potential, int on_device) {
    try {
        if (!t || !potential) throw Error(MSX_ERR_INVALID, "null argument");
        if (on_device != 0 && on_device != 1) throw Error(MSX_ERR_INVALID, "on_device is 0 or 1");
        const size_t n = t->nG() * t->nz;
        MSX_HIP(hipMemcpyAsync(t->V.p, potential, sizeof(cx<double>) * n,
                               on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice,
                               t->stream));
        t->prepare();
        return MSX_OK;
    } catch (const Error &e) {
        return fail(e);
    }
}

int msx_table_eig_phases(const msx_table *t, double *ms) {
    try {
        if (!t || !ms) throw Error(MSX_ERR_INVALID, "null argument");
        long long tk[8];
        t->ticks.download(tk, 8, t->stream);
        MSX_HIP(hipStreamSynchronize(t->stream));
        int rate_khz = 0;
        MSX_HIP(hipDeviceGetAttribute(&rate_khz, hipDeviceAttributeWallClockRate, 0));
        for (int i = 0; i < 4; ++i) ms[i] = rate_khz > 0 ? (double)(tk[i + 1] - tk[i]) / rate_khz : 0.0;
        ms[4] = rate_khz > 0 ? (double)tk[5] / rate_khz : 0.0;
        ms[5] = (double)tk[6];
        ms[6] = (double)tk[7];
        return MSX_OK;
    } catch (const Error &e) {
        return fail(e);
    }
}

int msx_symmetric_eigensystem(int n, const double *A, int method, double *lam, double *Q) {
    try {
        if (!A || !lam || !Q) throw Error(MSX_ERR_INVALID, "null argument");
        if (n < 1 || n > 2048) throw Error(MSX_ERR_DIMS, "1 <= n <= 2048 required");
        if (method < 0 || method > 2) throw Error(MSX_ERR_INVALID, "method must be 0, 1 or 2");
        if (n == 1) {
            lam[0] = A[0];
            Q[0] = 1.0;
            return MSX_OK;
        }
        hipStream_t s;
        MSX_HIP(hipStreamCreate(&s));
        struct StreamGuard {
            hipStream_t s;
            ~StreamGuard() { (void)hipStreamDestroy(s); }
        } guard{s};
        const size_t n2 = (size_t)n * n;
        DBuf<double> a(n2), vm(n2), dcw, l(n), q(n2), qt(n2);
        DBuf<int> info(1);
        a.upload(A, n2, s);
        run_eigensolver(n, a.p, vm.p, dcw, l.p, q.p, qt.p, info.p, nullptr, s, method);
        int h_info = 0;
        info.download(&h_info, 1, s);
        l.download(lam, n, s);
        q.download(Q, n2, s);
        MSX_HIP(hipStreamSynchronize(s));
        if (h_info != 0) throw Error(MSX_ERR_EIGEN, "eigensolver did not converge");
        return MSX_OK;
    } catch (const Error &e) {
        return fail(e);
    }
}

int msx_table_destroy(msx_table *t) {
    delete t;
    return MSX_OK;
}

int msx_table_eigensystem(const msx_table *t, double *lam, double *Q) {
    try {
        if (!t || !lam || !Q) throw Error(MSX_ERR_INVALID, "null argument");
        t->lam.download(lam, t->nz, t->stream);
        t->Q.download(Q, (size_t)t->nz * t->nz, t->stream);
        MSX_HIP(hipStreamSynchronize(t->stream));
        return MSX_OK;
    } catch (const Error &e) {
        return fail(e);
    }
}

int msx_batch_create(msx_table *t, int n_cond, const int *n_channels, const int *chx,
                     const int *chy, const double *e, const double *wave_b, const double *wave_c,
                     const double *wave_a0, msx_batch **out) {
    try {
        if (!t || !out || !n_channels || !chx || !chy || !e || !wave_b || !wave_c || !wave_a0)
            throw Error(MSX_ERR_INVALID, "null argument");
        *out = nullptr;
        if (n_cond <= 0) throw Error(MSX_ERR_DIMS, "no conditions");
        std::unique_ptr<msx_batch> B(new msx_batch);
        B->t = t;
        B->n_inst = n_cond;
        B->nc.assign(n_channels, n_channels + n_cond);
        B->off.resize(n_cond);
        B->spec.resize(n_cond);
        long long tot = 0;
        const size_t nG = t->nG();
        std::vector<int> chmap((size_t)n_cond * nG, -1);
        for (int b = 0; b < n_cond; ++b) {
            if (B->nc[b] <= 0) throw Error(MSX_ERR_DIMS, "condition without channels");
            if ((size_t)B->nc[b] > nG) throw Error(MSX_ERR_INVALID, "more channels than table entries");
            B->off[b] = tot;
            B->max_nc = std::max(B->max_nc, B->nc[b]);
            int sp = -1;
            for (int j = 0; j < B->nc[b]; ++j) {
                const int x = chx[tot + j], y = chy[tot + j];
                if (x < 0 || x >= t->nkx || y < 0 || y >= t->nky)
                    throw Error(MSX_ERR_INVALID, "channel index outside the table");
                int &slot = chmap[(size_t)b * nG + (size_t)x * t->nky + y];
                if (slot >= 0) throw Error(MSX_ERR_INVALID, "duplicate channel in a condition");
                slot = j;
                if (sp < 0 && x == 0 && y == 0) sp = j;
            }
            if (sp < 0) throw Error(MSX_ERR_NO_SPECULAR, "specular channel missing");
            B->spec[b] = sp;
            tot += B->nc[b];
        }
        B->total_ch = tot;
        hipStream_t s = t->stream;
        B->nc_d.alloc(n_cond);
        B->nc_d.upload(B->nc.data(), n_cond, s);
        B->off_d.alloc(n_cond);
        B->off_d.upload(B->off.data(), n_cond, s);
        B->spec_d.alloc(n_cond);
        B->spec_d.upload(B->spec.data(), n_cond, s);
        B->chx_d.alloc(tot);
        B->chx_d.upload(chx, tot, s);
        B->chy_d.alloc(tot);
        B->chy_d.upload(chy, tot, s);
        B->chmap_d.alloc(chmap.size());
        B->chmap_d.upload(chmap.data(), chmap.size(), s);
        B->e_d.alloc(tot);
        B->e_d.upload(e, tot, s);
        B->c_d.alloc(tot);
        B->c_d.upload(reinterpret_cast<const cx<double> *>(wave_c), tot, s);
        B->b_d.alloc(tot);
        B->b_d.upload(reinterpret_cast<const cx<double> *>(wave_b), tot, s);
        B->a0_d.alloc(n_cond);
        B->a0_d.upload(reinterpret_cast<const cx<double> *>(wave_a0), n_cond, s);
        std::vector<cx<double>> bsrc(n_cond);
        for (int b = 0; b < n_cond; ++b) {
            const long long g = B->off[b] + B->spec[b];
            bsrc[b] = mk(wave_b[2 * g], wave_b[2 * g + 1]);
        }
        B->bsrc_d.alloc(n_cond);
        B->bsrc_d.upload(bsrc.data(), n_cond, s);
        B->cf_d.alloc(tot);
        {
            std::vector<cx<float>> cf(tot);
            for (long long i = 0; i < tot; ++i)
                cf[i] = mk((float)wave_c[2 * i], (float)wave_c[2 * i + 1]);
            B->cf_d.upload(cf.data(), tot, s);
            MSX_HIP(hipStreamSynchronize(s));
        }
        B->fac_d.alloc((size_t)tot * t->nz);
        B->facf_d.alloc((size_t)tot * t->nz);
        B->psi_d.alloc((size_t)tot * t->nz);
        MSX_HIP(hipMemsetAsync(B->psi_d.p, 0, sizeof(cx<double>) * (size_t)tot * t->nz, t->stream));
        B->P_d.alloc(tot);
        MSX_HIP(hipStreamSynchronize(s));
        *out = B.release();
        return MSX_OK;
    } catch (const Error &e) {
        return fail(e);
    }
}

int msx_batch_destroy(msx_batch *b) {
    delete b;
    return MSX_OK;
}

int msx_batch_solve(msx_batch *B, const msx_solve_options *opt) {
    try {
        if (!B) throw Error(MSX_ERR_INVALID, "null batch");
        check_options(opt);
        msx_table &t = *B->t;
        B->build_factors();  // the table's eigensystem may have been re-prepared
        msx_batch::need(B->src, (size_t)B->n_elem());
        const Layout L = B->layout();
            // the source vector is read only by the Gauss-Seidel right-hand side; block Jacobi
        // forms D^-1 of the scattering source from the channel tables
        const cx<double> *src = nullptr;
        if (opt->method == MSX_METHOD_GAUSS_SEIDEL) {
            msx_batch::need(B->source, (size_t)B->n_elem());
            launch_set_source(L, B->spec_d.p, B->bsrc_d.p, B->source.p, t.stream);
            src = B->source.p;
        }
        solve_system(*B, *opt, src, B->psi_d.p, false, B->iters, B->info, B->resid, true);
        Layout Lpsi = L;
        Lpsi.nm = B->psi_nm ? 1 : 0;
        launch_intensities(Lpsi, B->psi_d.p, B->b_d.p, B->a0_d.p, B->spec_d.p, B->P_d.p, t.stream);
        MSX_HIP(hipStreamSynchronize(t.stream));
        B->solved = true;
        int missed = 0;
        for (int b = 0; b < B->n_inst; ++b) missed += B->info[b] ? 1 : 0;
        if (missed) {
            set_last_error(std::to_string(missed) + " of " + std::to_string(B->n_inst) +
                           " conditions did not reach the tolerance within maxiter");
            return MSX_ERR_NOT_CONVERGED;
        }
        set_last_error("");
        return MSX_OK;
    } catch (const Error &e) {
        return fail(e);
    }
}

int msx_batch_set_state(msx_batch *B, const double *psi) {
    try {
        if (!B || !psi) throw Error(MSX_ERR_INVALID, "null argument");
        B->psi_d.upload(reinterpret_cast<const cx<double> *>(psi), (size_t)B->n_elem(), B->t->stream);
        B->psi_nm = false;
        MSX_HIP(hipStreamSynchronize(B->t->stream));
        return MSX_OK;
    } catch (const Error &e) {
        return fail(e);
    }
}

int msx_batch_state(const msx_batch *B, double *psi) {
    try {
        if (!B || !psi) throw Error(MSX_ERR_INVALID, "null argument");
        // the order the state is kept in is internal: callers always see the batch layout
        const_cast<msx_batch *>(B)->psi_channel_major();
        B->psi_d.download(reinterpret_cast<cx<double> *>(psi), (size_t)B->n_elem(), B->t->stream);
        MSX_HIP(hipStreamSynchronize(B->t->stream));
        return MSX_OK;
    } catch (const Error &e) {
        return fail(e);
    }
}

int msx_batch_apply_lower(msx_batch *B, int inverse, const double *state_in,
                          double *state_out) {
    try {
        if (!B || !state_in || !state_out) throw Error(MSX_ERR_INVALID, "null argument");
        if (inverse != 0 && inverse != 1) throw Error(MSX_ERR_INVALID, "inverse is 0 or 1");
        msx_table &t = *B->t;
        hipStream_t s = t.stream;
        const size_t n = (size_t)B->n_elem();
        const Layout L = B->layout();
        B->build_factors();  // the table's eigensystem may have been re-prepared
        msx_batch::need(B->tmp1, n);
        msx_batch::need(B->tmp2, n);
        msx_batch::need(B->tmp3, n);
        B->tmp1.upload(reinterpret_cast<const cx<double> *>(state_in), n, s);
        cx<double> *res = B->tmp1.p;
        if (inverse) {
            // L^-1: the block forward substitution of the Gauss-Seidel split, in place
            launch_lower_solve(L, t.V.p, t.Q.p, t.QT.p, t.lam.p, B->e_d.p, B->c_d.p, B->fac_d.p,
                               B->tmp1.p, nullptr, nullptr, 0.0, nullptr, s);
        } else {
            // L x = D x + sum_{i<j} V[G_j - G_i] x_i
            launch_coupling<double>(L, t.V.p, B->tmp1.p, B->tmp2.p, COUPLE_LOWER, nullptr, s);
            launch_apply_diag(L, t.Q.p, t.QT.p, t.lam.p, B->e_d.p, B->c_d.p, B->tmp1.p,
                              B->tmp3.p, nullptr, s, B->tmp2.p);
            res = B->tmp3.p;
        }
        MSX_HIP(hipMemcpyAsync(state_out, res, sizeof(cx<double>) * n, hipMemcpyDeviceToHost, s));
        MSX_HIP(hipStreamSynchronize(s));
        return MSX_OK;
    } catch (const Error &e) {
        return fail(e);
    }
}

int msx_batch_tuning(const msx_batch *B, char *buf, int buflen) {
    if (!B || !buf || buflen <= 0) {
        set_last_error("null argument");
        return MSX_ERR_INVALID;
    }
    const std::string &r = B->tuning.done ? B->tuning.report : std::string("{}");
    std::strncpy(buf, r.c_str(), (size_t)buflen - 1);
    buf[buflen - 1] = 0;
    return (int)r.size() < buflen ? MSX_OK : MSX_ERR_INVALID;
}

int msx_batch_intensities(const msx_batch *B, double *P) {
    try {
        if (!B || !P) throw Error(MSX_ERR_INVALID, "null argument");
        B->P_d.download(P, (size_t)B->total_ch, B->t->stream);
        MSX_HIP(hipStreamSynchronize(B->t->stream));
        return MSX_OK;
    } catch (const Error &e) {
        return fail(e);
    }
}

int msx_batch_stats(const msx_batch *B, int *iterations, int *info, double *residual) {
    if (!B) {
        set_last_error("null batch");
        return MSX_ERR_INVALID;
    }
    for (int b = 0; b < B->n_inst; ++b) {
        if (iterations) iterations[b] = b < (int)B->iters.size() ? B->iters[b] : 0;
        if (info) info[b] = b < (int)B->info.size() ? B->info[b] : 0;
        if (residual) residual[b] = b < (int)B->resid.size() ? B->resid[b] : 0.0;
    }
    return MSX_OK;
}

int msx_batch_device_state(const msx_batch *B, void **psi_device) {
    if (!B || !psi_device) {
        set_last_error("null argument");
        return MSX_ERR_INVALID;
    }
    try {
        const_cast<msx_batch *>(B)->psi_channel_major();
        MSX_HIP(hipStreamSynchronize(B->t->stream));
    } catch (const Error &e) {
        return fail(e);
    }
    *psi_device = B->psi_d.p;
    return MSX_OK;
}

int msx_batch_total_channels(const msx_batch *B, long long *total) {
    if (!B || !total) {
        set_last_error("null argument");
        return MSX_ERR_INVALID;
    }
    *total = B->total_ch;
    return MSX_OK;
}

int msx_batch_adjoint(msx_batch *B, const msx_solve_options *opt, const double *dL_dP,
                      double *grad_potential) {
    try {
        if (!B || !dL_dP || !grad_potential) throw Error(MSX_ERR_INVALID, "null argument");
        if (!B->solved) throw Error(MSX_ERR_INVALID, "adjoint before a forward solve");
        check_options(opt);
        msx_table &t = *B->t;
        hipStream_t s = t.stream;
        const Layout L = B->layout();
        const size_t n = (size_t)B->n_elem();
        B->psi_channel_major();
        msx_batch::need(B->dLdP_d, (size_t)B->total_ch);
        B->dLdP_d.upload(dL_dP, (size_t)B->total_ch, s);
        // buffers owned by the batch, reused by later adjoints
        msx_batch::need(B->adj_src, n);
        msx_batch::need(B->adj_state, n);
        launch_adjoint_source(L, B->spec_d.p, B->psi_d.p, B->b_d.p, B->a0_d.p, B->dLdP_d.p,
                              B->adj_src.p, s);
        std::vector<int> it, inf;
        std::vector<double> res;
        solve_system(*B, *opt, B->adj_src.p, B->adj_state.p, true, it, inf, res, false);
        const size_t nG = t.nG();
        msx_batch::need(B->grad_d, nG * t.nz);
        MSX_HIP(hipMemsetAsync(B->grad_d.p, 0, sizeof(cx<double>) * nG * t.nz, s));
        const bool dft_form = opt->coupling == 0 && (t.dft.np == 16 || t.dft.np == 32);
        {
            // correlation on the grid: two forward 2-D transforms per condition and node (four
            // complex GEMMs), one inverse per node; the direct form gathers nc^2 products
            Work fl, by;
            if (profiling()) {
                const double nz = t.nz, nkx = t.nkx, nky = t.nky, ng = nkx * nky;
                double flops = 0, bytes = 0;
                for (int b = 0; b < B->n_inst; ++b) {
                    flops += dft_form ? nz * (16.0 * ng * (nkx + nky) + 8.0 * ng)
                                      : 8.0 * B->nc[b] * (double)B->nc[b] * nz;
                    bytes += 2.0 * B->nc[b] * nz * sizeof(cx<double>);
                }
                fl = Work(flops);
                by = Work(bytes + ng * nz * sizeof(cx<double>));
            }
            ProfScope ps(dft_form ? "gradient_dft_f64" : "gradient_direct_f64", s, fl, by);
            if (dft_form)
                launch_dft_gradient(L, t.dft.d64(), B->adj_state.p, B->psi_d.p, B->grad_d.p,
                                    B->grad_part, s);
            else
                launch_potential_gradient_map(L, B->chmap_d.p, B->adj_state.p, B->psi_d.p,
                                              B->grad_d.p, s);
        }
        B->grad_d.download(reinterpret_cast<cx<double> *>(grad_potential), nG * t.nz, s);
        MSX_HIP(hipStreamSynchronize(s));
        for (int b = 0; b < B->n_inst; ++b)
            if (inf[b]) return MSX_ERR_NOT_CONVERGED;
        return MSX_OK;
    } catch (const Error &e) {
        return fail(e);
    }
}

// ---- drop-in group ----------------------------------------------------------------------

int msx_run_multiscat_fortran_ex(int flag, int nsig, int nkx, int nky, int nz, const double *V,
                                 const double *e, const double *wave_a, const double *wave_b,
                                 const double *wave_c, const double *T, double *out,
                                 int *iterations, double *residual) {
    try {
        if (nkx <= 0 || nky <= 0 || nz <= 0) return MSX_ERR_DIMS;  // bindings.f90:42-45
        if (!V || !e || !wave_a || !wave_b || !wave_c || !T || !out)
            throw Error(MSX_ERR_INVALID, "null argument");
        std::vector<cx<double>> Vr;
        table_from_fortran(nkx, nky, nz, V, Vr);
        std::vector<double> Tr;
        transpose_square(nz, T, Tr);
        TableGuard tg;
        int rc = msx_table_create(nkx, nky, nz, reinterpret_cast<const double *>(Vr.data()),
                                  Tr.data(), &tg.t);
        if (rc != MSX_OK) return rc;
        BatchGuard bg;
        bg.b = full_grid_batch(tg.t, e, wave_b, wave_c, wave_a);
        msx_solve_options o;
        msx_default_options(&o, nsig);
        o.neumann = flag == 1 ? 1 : 0;  // multiscat_gmres.f90:469
        rc = msx_batch_solve(bg.b, &o);
        if (rc != MSX_OK && rc != MSX_ERR_NOT_CONVERGED) return rc;
        const size_t nG = (size_t)nkx * nky;
        std::vector<cx<double>> psi(nG * nz);
        int r2 = msx_batch_state(bg.b, reinterpret_cast<double *>(psi.data()));
        if (r2 != MSX_OK) return r2;
        // (nkx, nky, nz) column-major output (multiscat.f90:53-59)
        for (int k = 0; k < nz; ++k)
            for (int iy = 0; iy < nky; ++iy)
                for (int ix = 0; ix < nkx; ++ix) {
                    const size_t f = (size_t)ix + (size_t)nkx * ((size_t)iy + (size_t)nky * k);
                    const cx<double> v = psi[((size_t)ix * nky + iy) * nz + k];
                    out[2 * f] = v.re;
                    out[2 * f + 1] = v.im;
                }
        if (iterations) *iterations = bg.b->iters.empty() ? 0 : bg.b->iters[0];
        if (residual) *residual = bg.b->resid.empty() ? 0.0 : bg.b->resid[0];
        return rc == MSX_ERR_NOT_CONVERGED ? MSX_ERR_NOT_CONVERGED : MSX_OK;
    } catch (const Error &e) {
        return fail(e);
    }
}

int msx_run_multiscat_fortran(int flag, int nsig, int nkx, int nky, int nz, const double *V,
                              const double *e, const double *wave_a, const double *wave_b,
                              const double *wave_c, const double *T, double *out) {
    return msx_run_multiscat_fortran_ex(flag, nsig, nkx, nky, nz, V, e, wave_a, wave_b, wave_c, T,
                                        out, nullptr, nullptr);
}

int msx_debug_build_preconditioner(int nkx, int nky, int nz, const double *V, const double *e,
                                   const double *T, double *eigenvalues, double *factors,
                                   double *eigenvectors) {
    try {
        if (nkx <= 0 || nky <= 0 || nz <= 0) return MSX_ERR_DIMS;
        if (!V || !e || !T || !eigenvalues || !factors || !eigenvectors)
            throw Error(MSX_ERR_INVALID, "null argument");
        std::vector<cx<double>> Vr;
        table_from_fortran(nkx, nky, nz, V, Vr);
        std::vector<double> Tr;
        transpose_square(nz, T, Tr);
        TableGuard tg;
        int rc = msx_table_create(nkx, nky, nz, reinterpret_cast<const double *>(Vr.data()),
                                  Tr.data(), &tg.t);
        if (rc != MSX_OK) return rc;
        BatchGuard bg;
        bg.b = full_grid_batch(tg.t, e, nullptr, nullptr, nullptr);
        bg.b->build_factors();
        hipStream_t s = tg.t->stream;
        tg.t->lam.download(eigenvalues, nz, s);
        // (nz, nz) column-major: element (z, m) = Q[z][m] = QT[m][z] row-major
        tg.t->QT.download(eigenvectors, (size_t)nz * nz, s);
        // (nz, nkx*nky) column-major: element (z, j) = fac[j][z]
        bg.b->fac_d.download(factors, (size_t)nz * nkx * nky, s);
        MSX_HIP(hipStreamSynchronize(s));
        return MSX_OK;
    } catch (const Error &e) {
        return fail(e);
    }
}

int msx_debug_apply_upper_block(int nkx, int nky, int nz, const double *V, const double *x,
                                double *y) {
    try {
        if (nkx <= 0 || nky <= 0 || nz <= 0) return MSX_ERR_DIMS;
        if (!V || !x || !y) throw Error(MSX_ERR_INVALID, "null argument");
        std::vector<cx<double>> Vr;
        table_from_fortran(nkx, nky, nz, V, Vr);
        std::vector<double> Tr((size_t)nz * nz, 0.0);
        for (int i = 0; i < nz; ++i) Tr[(size_t)i * nz + i] = 1.0;
        TableGuard tg;
        int rc = msx_table_create(nkx, nky, nz, reinterpret_cast<const double *>(Vr.data()),
                                  Tr.data(), &tg.t);
        if (rc != MSX_OK) return rc;
        BatchGuard bg;
        bg.b = full_grid_batch(tg.t, nullptr, nullptr, nullptr, nullptr);
        const size_t n = (size_t)nkx * nky * nz;
        DBuf<cx<double>> xin(n), yout(n);
        hipStream_t s = tg.t->stream;
        xin.upload(reinterpret_cast<const cx<double> *>(x), n, s);
        launch_coupling<double>(bg.b->layout(), tg.t->V.p, xin.p, yout.p, COUPLE_UPPER, nullptr, s);
        yout.download(reinterpret_cast<cx<double> *>(y), n, s);
        MSX_HIP(hipStreamSynchronize(s));
        return MSX_OK;
    } catch (const Error &e) {
        return fail(e);
    }
}

int msx_debug_apply_coupling(int nkx, int nky, int nz, const double *V, int nc, const int *chx,
                             const int *chy, const double *x, double *y, int mode,
                             int precision, int algorithm) {
    try {
        if (nkx <= 0 || nky <= 0 || nz <= 0 || nc <= 0) return MSX_ERR_DIMS;
        if (!V || !chx || !chy || !x || !y) throw Error(MSX_ERR_INVALID, "null argument");
        if (mode < 0 || mode > 2 || precision < 0 || precision > 1 || algorithm < 0 || algorithm > 5)
            throw Error(MSX_ERR_INVALID, "bad mode/precision/algorithm");
        if (algorithm == 5 && (precision != 1 || mode != COUPLE_FULL))
            throw Error(MSX_ERR_INVALID, "algorithm 5 (split FP16) is the FP32 full coupling only");
        const bool split = algorithm == 5;
        // DFT form with the built-in node grouping (1), or 1, 2 or 8 nodes per wavefront
        const int npw = algorithm == 2 ? 1 : algorithm == 3 ? 2 : algorithm == 4 ? 8 : 0;
        if (algorithm > 1) algorithm = 1;
        std::vector<double> Tr((size_t)nz * nz, 0.0);
        for (int i = 0; i < nz; ++i) Tr[(size_t)i * nz + i] = 1.0;
        TableGuard tg;
        int rc = msx_table_create(nkx, nky, nz, V, Tr.data(), &tg.t);
        if (rc != MSX_OK) return rc;
        // the specular channel is not needed for a bare coupling: build the layout directly
        msx_table &t = *tg.t;
        hipStream_t s = t.stream;
        DBuf<int> nc_d(1), chx_d(nc), chy_d(nc);
        DBuf<long long> off_d(1);
        long long zero = 0;
        nc_d.upload(&nc, 1, s);
        off_d.upload(&zero, 1, s);
        chx_d.upload(chx, nc, s);
        chy_d.upload(chy, nc, s);
        const Layout L{nkx, nky, nz, 1, nc, nc_d.p, off_d.p, chx_d.p, chy_d.p};
        const size_t n = (size_t)nc * nz;
        DBuf<cx<double>> xin(n), yout(n);
        xin.upload(reinterpret_cast<const cx<double> *>(x), n, s);
        if (algorithm == 1 && mode != COUPLE_FULL)
            throw Error(MSX_ERR_INVALID, "the DFT form exists for the full coupling only");
        if (algorithm == 1 && t.dft.np == 0) throw Error(MSX_ERR_INVALID, "table too large for the DFT form");
        if (precision == 0) {
            if (algorithm == 0) launch_coupling<double>(L, t.V.p, xin.p, yout.p, mode, nullptr, s);
            else launch_dft_coupling<double>(L, t.dft.d64(), xin.p, yout.p, nullptr, s, -1, nullptr,
                                             nullptr, npw);
        } else if (split) {
            // the split-FP16 kernel takes node-major vectors: transposed here
            std::vector<cx<float>> hx(n);
            for (int j = 0; j < nc; ++j)
                for (int z = 0; z < nz; ++z) {
                    const size_t cm = (size_t)j * nz + z;
                    hx[(size_t)z * nc + j] = mk((float)x[2 * cm], (float)x[2 * cm + 1]);
                }
            DBuf<cx<float>> xf(n), yf(n);
            xf.upload(hx.data(), n, s);
            Layout Ln = L;
            Ln.nm = 1;
            Ln.total = nc;
            if (!dft_h3_fits(Ln, t.dft.d32()))
                throw Error(MSX_ERR_INVALID, "split-FP16 DFT coupling: square tables up to 32 x 32");
            launch_dft_coupling_h3(Ln, t.dft.d32(), xf.p, yf.p, nullptr, s);
            yf.download(hx.data(), n, s);
            MSX_HIP(hipStreamSynchronize(s));
            for (int j = 0; j < nc; ++j)
                for (int z = 0; z < nz; ++z) {
                    const size_t cm = (size_t)j * nz + z;
                    y[2 * cm] = hx[(size_t)z * nc + j].re;
                    y[2 * cm + 1] = hx[(size_t)z * nc + j].im;
                }
            return MSX_OK;
        } else {
            DBuf<cx<float>> xf(n), yf(n);
            launch_to_float((long long)n, xin.p, xf.p, s);
            if (algorithm == 0) launch_coupling<float>(L, t.Vf.p, xf.p, yf.p, mode, nullptr, s);
            else launch_dft_coupling<float>(L, t.dft.d32(), xf.p, yf.p, nullptr, s, -1, nullptr,
                                            nullptr, npw);
            std::vector<cx<float>> hf(n);
            yf.download(hf.data(), n, s);
            MSX_HIP(hipStreamSynchronize(s));
            for (size_t i = 0; i < n; ++i) {
                y[2 * i] = hf[i].re;
                y[2 * i + 1] = hf[i].im;
            }
            return MSX_OK;
        }
        yout.download(reinterpret_cast<cx<double> *>(y), n, s);
        MSX_HIP(hipStreamSynchronize(s));
        return MSX_OK;
    } catch (const Error &e) {
        return fail(e);
    }
}

int msx_debug_apply_diagonal(int nz, const double *potential00, const double *kinetic, int nc,
                             const double *e, const double *wave_c, const double *state_in,
                             double *state_out, int inverse, int precision, int variant) {
    try {
        if (nz <= 0 || nc <= 0) return MSX_ERR_DIMS;
        if (!potential00 || !kinetic || !e || !wave_c || !state_in || !state_out)
            throw Error(MSX_ERR_INVALID, "null argument");
        if (precision < 0 || precision > 1 || variant < 0 || variant > 4 || inverse < 0 ||
            inverse > 1)
            throw Error(MSX_ERR_INVALID, "bad inverse/precision/variant");
        // variants 3 and 4: split FP16 with 16 and 8 channels per wavefront
        const int h3_cpw = variant == 4 ? 8 : 16;
        if (variant == 4) variant = 3;
        if (variant == 3 && (precision != 1 || inverse != 1))
            throw Error(MSX_ERR_INVALID, "variants 3, 4 (split FP16) are the FP32 inverse only");
        if (variant == 3 && h3_padded_size(nz) == 0)
            throw Error(MSX_ERR_INVALID, "split-FP16 diagonal blocks need nz <= 224");
        // variant 2: the matrix-core form with 8 channels per wavefront (FP64)
        const int cpw = variant == 2 ? 8 : 0;
        if (variant == 2) variant = 0;
        if (variant == 0 && !diag_mfma_supported(nz))
            throw Error(MSX_ERR_INVALID, "matrix-core diagonal blocks need nz <= 320");
        TableGuard tg;
        int rc = msx_table_create(1, 1, nz, potential00, kinetic, &tg.t);
        if (rc != MSX_OK) return rc;
        msx_table &t = *tg.t;
        hipStream_t s = t.stream;
        // channels need no grid position here: a layout of nc rows, all of instance 0
        std::vector<int> zeros(nc, 0);
        DBuf<int> nc_d(1), ch_d(nc);
        DBuf<long long> off_d(1);
        long long zero = 0;
        nc_d.upload(&nc, 1, s);
        off_d.upload(&zero, 1, s);
        ch_d.upload(zeros.data(), nc, s);
        const Layout L{1, 1, nz, 1, nc, nc_d.p, off_d.p, ch_d.p, ch_d.p};
        const size_t n = (size_t)nc * nz;
        DBuf<double> e_d(nc), fac(n);
        DBuf<float> facf(n);
        DBuf<cx<double>> c_d(nc), xin(n), yout(n);
        e_d.upload(e, nc, s);
        c_d.upload(reinterpret_cast<const cx<double> *>(wave_c), nc, s);
        xin.upload(reinterpret_cast<const cx<double> *>(state_in), n, s);
        launch_preconditioner_factors_qt(nc, nz, t.Q.p, t.QT.p, t.lam.p, e_d.p, fac.p, s);
        if (variant == 3) {
            // the split-FP16 kernel takes node-major vectors and factors: transposed here
            std::vector<double> hfac(n);
            fac.download(hfac.data(), n, s);
            MSX_HIP(hipStreamSynchronize(s));
            std::vector<cx<float>> hx(n);
            std::vector<float> hf(n);
            for (int j = 0; j < nc; ++j)
                for (int z = 0; z < nz; ++z) {
                    const size_t cm = (size_t)j * nz + z, nm = (size_t)z * nc + j;
                    hx[nm] = mk((float)state_in[2 * cm], (float)state_in[2 * cm + 1]);
                    hf[nm] = (float)hfac[cm];
                }
            DBuf<cx<float>> cf(nc), xf(n), yf(n);
            DBuf<float> facn(n);
            launch_to_float(nc, c_d.p, cf.p, s);
            xf.upload(hx.data(), n, s);
            facn.upload(hf.data(), n, s);
            Layout Ln = L;
            Ln.nm = 1;
            Ln.total = nc;
            if (!diag_h3_fits(Ln, t.h3()))
                throw Error(MSX_ERR_INVALID, "split-FP16 diagonal blocks: shape outside the kernel");
            launch_diag_h3(Ln, t.h3(), t.lam.p, e_d.p, cf.p, facn.p, xf.p, nullptr, yf.p, 1.0f,
                           nullptr, s, h3_cpw);
            yf.download(hx.data(), n, s);
            MSX_HIP(hipStreamSynchronize(s));
            for (int j = 0; j < nc; ++j)
                for (int z = 0; z < nz; ++z) {
                    const size_t cm = (size_t)j * nz + z, nm = (size_t)z * nc + j;
                    state_out[2 * cm] = hx[nm].re;
                    state_out[2 * cm + 1] = hx[nm].im;
                }
            return MSX_OK;
        }
        if (precision == 0) {
            if (variant == 0)
                launch_diag_blocks<double>(L, t.eig64(), e_d.p, c_d.p, fac.p, xin.p, nullptr,
                                           yout.p, 1.0, inverse, nullptr, s, cpw);
            else if (inverse)
                launch_dinv<double>(L, t.Q.p, t.QT.p, t.lam.p, e_d.p, c_d.p, fac.p, xin.p,
                                    nullptr, yout.p, 1.0, nullptr, s);
            else
                launch_apply_diag(L, t.Q.p, t.QT.p, t.lam.p, e_d.p, c_d.p, xin.p, yout.p,
                                  nullptr, s);
        } else {
            DBuf<cx<float>> cf(nc), xf(n), yf(n);
            launch_to_float(nc, c_d.p, cf.p, s);
            launch_to_float((long long)n, xin.p, xf.p, s);
            launch_to_float_real((long long)n, fac.p, facf.p, s);
            if (variant == 0)
                launch_diag_blocks<float>(L, t.eig32(), e_d.p, cf.p, facf.p, xf.p, nullptr, yf.p,
                                          1.0f, inverse, nullptr, s);
            else if (inverse)
                launch_dinv<float>(L, t.Qf.p, t.QTf.p, t.lam.p, e_d.p, cf.p, facf.p, xf.p, nullptr,
                                   yf.p, 1.0f, nullptr, s);
            else
                throw Error(MSX_ERR_INVALID, "forward FP32 blocks exist in the matrix-core form only");
            std::vector<cx<float>> h(n);
            yf.download(h.data(), n, s);
            MSX_HIP(hipStreamSynchronize(s));
            for (size_t i = 0; i < n; ++i) {
                state_out[2 * i] = h[i].re;
                state_out[2 * i + 1] = h[i].im;
            }
            return MSX_OK;
        }
        yout.download(reinterpret_cast<cx<double> *>(state_out), n, s);
        MSX_HIP(hipStreamSynchronize(s));
        return MSX_OK;
    } catch (const Error &e) {
        return fail(e);
    }
}

int msx_debug_solve_lower_block(int nkx, int nky, int nz, const double *V, const double *wave_c,
                                const double *e, const double *T, const double *x, double *y) {
    try {
        if (nkx <= 0 || nky <= 0 || nz <= 0) return MSX_ERR_DIMS;
        if (!V || !wave_c || !e || !T || !x || !y) throw Error(MSX_ERR_INVALID, "null argument");
        std::vector<cx<double>> Vr;
        table_from_fortran(nkx, nky, nz, V, Vr);
        std::vector<double> Tr;
        transpose_square(nz, T, Tr);
        TableGuard tg;
        int rc = msx_table_create(nkx, nky, nz, reinterpret_cast<const double *>(Vr.data()),
                                  Tr.data(), &tg.t);
        if (rc != MSX_OK) return rc;
        BatchGuard bg;
        bg.b = full_grid_batch(tg.t, e, nullptr, wave_c, nullptr);
        bg.b->build_factors();
        const size_t n = (size_t)nkx * nky * nz;
        DBuf<cx<double>> xin(n);
        hipStream_t s = tg.t->stream;
        xin.upload(reinterpret_cast<const cx<double> *>(x), n, s);
        launch_lower_solve(bg.b->layout(), tg.t->V.p, tg.t->Q.p, tg.t->QT.p, tg.t->lam.p,
                           bg.b->e_d.p, bg.b->c_d.p, bg.b->fac_d.p, xin.p, nullptr, nullptr, 0.0,
                           nullptr, s);
        xin.download(reinterpret_cast<cx<double> *>(y), n, s);
        MSX_HIP(hipStreamSynchronize(s));
        return MSX_OK;
    } catch (const Error &e) {
        return fail(e);
    }
}

}  // extern "C"
