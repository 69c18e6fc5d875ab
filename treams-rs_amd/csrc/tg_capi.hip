// C ABI of the cluster-coupling library (declarations and reference citations in
// include/treams_gpu.h).
#include <cmath>
#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "tg_gmres.h"
#include "tg_internal.h"
#include "tg_precond.h"

namespace tg {

static thread_local std::string g_error;

void set_error(const char *fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_error = buf;
}

// Released device blocks are kept in a process-wide cache and handed to later requests
// of the same size, else of at least half their size: hipFree synchronises the device
// and unmaps the pages, which costs milliseconds per large block, and every solve
// allocates its scratch again.  Work queued before a release may still use the block,
// so taking one from the cache synchronises the device first, unless the block was
// released clean (after the streams that used it were synchronised: a cluster or
// factor being destroyed).  Blocks above kCacheBlockMax are freed at once, the cache
// holds at most kCacheBytes, it is emptied when an allocation fails (and the
// allocation retried); when the last cluster is destroyed the blocks above
// kKeepBlockMax are freed, and the small ones (up to kKeepBytes) stay for the next
// cluster, so that creating and destroying small clusters allocates nothing.  Pinned
// host blocks (HostBuf) are pooled the same way, up to kPinnedKeep blocks.
namespace {
constexpr size_t kCacheBytes = (size_t)32 << 30;
constexpr size_t kCacheBlockMax = (size_t)8 << 30;
constexpr size_t kKeepBlockMax = (size_t)256 << 20;
constexpr size_t kKeepBytes = (size_t)1 << 30;
constexpr size_t kPinnedKeep = 16;
std::mutex g_cache_mu;
struct CachedBlock {
    void *p;
    bool clean;
};
std::multimap<size_t, CachedBlock> g_cache;  // block size -> block
size_t g_cached = 0;
int g_live_clusters = 0;
std::multimap<size_t, void *> g_pinned;  // pinned host blocks by size
thread_local bool t_clean_release = false;

void *cache_take(size_t need, size_t &got, bool &clean) {
    std::lock_guard<std::mutex> lock(g_cache_mu);
    // a block of exactly this size first (repeated solves ask for the same sizes, and
    // taking a larger block would push its own request to a fresh allocation)
    auto it = g_cache.find(need);
    if (it == g_cache.end()) it = g_cache.lower_bound(need);
    if (it == g_cache.end() || it->first / 2 > need) return nullptr;
    void *p = it->second.p;
    clean = it->second.clean;
    got = it->first;
    g_cached -= got;
    g_cache.erase(it);
    return p;
}

bool cache_put(void *p, size_t bytes, bool clean) {
    std::lock_guard<std::mutex> lock(g_cache_mu);
    if (bytes > kCacheBlockMax || g_cached + bytes > kCacheBytes) return false;
    g_cache.emplace(bytes, CachedBlock{p, clean});
    g_cached += bytes;
    return true;
}

// free the cached blocks larger than `keep_block`, then the largest others until at
// most `keep_total` bytes remain
void cache_trim(size_t keep_block, size_t keep_total) {
    std::lock_guard<std::mutex> lock(g_cache_mu);
    for (auto it = g_cache.begin(); it != g_cache.end();) {
        if (it->first > keep_block) {
            (void)hipFree(it->second.p);
            g_cached -= it->first;
            it = g_cache.erase(it);
        } else {
            ++it;
        }
    }
    while (g_cached > keep_total && !g_cache.empty()) {
        auto it = std::prev(g_cache.end());
        (void)hipFree(it->second.p);
        g_cached -= it->first;
        g_cache.erase(it);
    }
}

void cache_clear() { cache_trim(0, 0); }

void cluster_born() {
    std::lock_guard<std::mutex> lock(g_cache_mu);
    ++g_live_clusters;
}

void cluster_died() {
    bool last;
    {
        std::lock_guard<std::mutex> lock(g_cache_mu);
        last = --g_live_clusters == 0;
    }
    if (last) cache_trim(kKeepBlockMax, kKeepBytes);
}

void *pinned_take(size_t need, size_t &got) {
    std::lock_guard<std::mutex> lock(g_cache_mu);
    auto it = g_pinned.lower_bound(need);
    if (it == g_pinned.end() || it->first / 2 > need) return nullptr;
    void *p = it->second;
    got = it->first;
    g_pinned.erase(it);
    return p;
}

bool pinned_put(void *p, size_t bytes) {
    std::lock_guard<std::mutex> lock(g_cache_mu);
    if (g_pinned.size() >= kPinnedKeep || bytes > ((size_t)16 << 20)) return false;
    g_pinned.emplace(bytes, p);
    return true;
}
}  // namespace

CleanReleases::CleanReleases() : prev(t_clean_release) { t_clean_release = true; }
CleanReleases::~CleanReleases() { t_clean_release = prev; }

int HostBuf::ensure(size_t need) {
    if (need <= bytes && ptr) return TG_OK;
    release();
    if (need == 0) return TG_OK;
    size_t got = 0;
    if (void *p = pinned_take(need, got)) {
        // a pinned block is released by its owner only after the stream that wrote it
        // was synchronised (the host has read it)
        ptr = p;
        bytes = got;
        return TG_OK;
    }
    hipError_t e = hipHostMalloc(&ptr, need, hipHostMallocDefault);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        ptr = nullptr;
        set_error("hipHostMalloc(%zu bytes): %s", need, hipGetErrorString(e));
        return e == hipErrorOutOfMemory ? TG_OUT_OF_MEMORY : TG_DEVICE_ERROR;
    }
    bytes = need;
    return TG_OK;
}

void HostBuf::release() {
    if (ptr && !pinned_put(ptr, bytes)) (void)hipHostFree(ptr);
    ptr = nullptr;
    bytes = 0;
}

int DevBuf::ensure(size_t need) {
    if (need <= bytes && ptr) return TG_OK;
    release();
    if (need == 0) return TG_OK;
    size_t got = 0;
    bool clean = false;
    if (void *p = cache_take(need, got, clean)) {
        hipError_t e = clean ? hipSuccess : hipDeviceSynchronize();
        if (e != hipSuccess) {
            (void)hipFree(p);
            set_error("hipDeviceSynchronize: %s", hipGetErrorString(e));
            return TG_DEVICE_ERROR;
        }
        ptr = p;
        bytes = got;
        return TG_OK;
    }
    hipError_t e = hipMalloc(&ptr, need);
    if (e == hipErrorOutOfMemory) {
        (void)hipGetLastError();
        cache_clear();
        e = hipMalloc(&ptr, need);
    }
    if (e != hipSuccess) {
        // a failed hipMalloc leaves its error as the runtime's last error, which the next
        // launch check (hipGetLastError) would report as its own: clear it here
        (void)hipGetLastError();
        ptr = nullptr;
        set_error("device allocation of %zu bytes failed: %s", need, hipGetErrorString(e));
        return e == hipErrorOutOfMemory ? TG_OUT_OF_MEMORY : TG_DEVICE_ERROR;
    }
    bytes = need;
    return TG_OK;
}

void DevBuf::release() {
    if (ptr && !cache_put(ptr, bytes, t_clean_release)) (void)hipFree(ptr);
    ptr = nullptr;
    bytes = 0;
}

int Timer::begin(const char *name, hipStream_t s, Rec &r) {
    if (!enabled) return TG_OK;
    // events are reused from collected records (creating two per region cost about as
    // much host time as recording them)
    for (hipEvent_t *e : {&r.a, &r.b}) {
        if (!spare.empty()) {
            *e = spare.back();
            spare.pop_back();
        } else {
            TG_HIP(hipEventCreate(e));
        }
    }
    TG_HIP(hipEventRecord(r.a, s));
    (void)name;
    return TG_OK;
}

int Timer::end(const char *name, hipStream_t s, Rec &r) {
    if (!enabled) return TG_OK;
    TG_HIP(hipEventRecord(r.b, s));
    pending[name].push_back(r);
    return TG_OK;
}

int Timer::collect() {
    for (auto &kv : pending) {
        auto &tot = totals[kv.first];
        for (auto &r : kv.second) {
            TG_HIP(hipEventSynchronize(r.b));
            float ms = 0.f;
            TG_HIP(hipEventElapsedTime(&ms, r.a, r.b));
            tot.first += ms;
            tot.second += 1;
            spare.push_back(r.a);
            spare.push_back(r.b);
        }
    }
    pending.clear();
    return TG_OK;
}

void Timer::destroy_events() {
    for (hipEvent_t e : spare) (void)hipEventDestroy(e);
    spare.clear();
}

// Upload a host plan into one device buffer and fill the device view.
static int upload_plan(const HostPlan &hp, DevBuf &buf, DevPlan &dp, std::vector<int> *ntiles = nullptr) {
    const size_t HH = (size_t)hp.H * hp.H;
    const size_t nterms = hp.w.size();
    // transposed plan: terms grouped by harmonic index, entry order kept
    std::vector<int32_t> tstarts(hp.table_len + 1, 0), tentry(nterms);
    std::vector<double> tw(nterms);
    for (size_t t = 0; t < nterms; ++t) tstarts[hp.hidx[t] + 1]++;
    for (int i = 0; i < hp.table_len; ++i) tstarts[i + 1] += tstarts[i];
    std::vector<int32_t> fill(tstarts.begin(), tstarts.end() - 1);
    for (size_t e = 0; e < HH; ++e)
        for (int32_t t = hp.starts[e]; t < hp.starts[e + 1]; ++t) {
            int32_t dst = fill[hp.hidx[t]]++;
            int cross = (t - hp.starts[e]) >= hp.nsame[e];
            tentry[dst] = (int32_t)e | (cross ? (int32_t)0x40000000 : 0);
            tw[dst] = hp.w[t];
        }
    // source-mode subsets per harmonic index and the subset column of every term
    const int T = hp.table_len;
    std::vector<std::vector<char>> used(T, std::vector<char>(hp.H, 0));
    for (size_t e = 0; e < HH; ++e)
        for (int32_t t = hp.starts[e]; t < hp.starts[e + 1]; ++t) used[hp.hidx[t]][e / hp.H] = 1;
    std::vector<int32_t> soff(T + 1, 0), smode, rank(T * hp.H, -1), ucol(nterms);
    for (int t = 0; t < T; ++t) {
        for (int a = 0; a < hp.H; ++a)
            if (used[t][a]) {
                rank[t * hp.H + a] = (int32_t)smode.size() - soff[t];
                smode.push_back(a);
            }
        soff[t + 1] = (int32_t)smode.size();
    }
    for (size_t e = 0; e < HH; ++e)
        for (int32_t t = hp.starts[e]; t < hp.starts[e + 1]; ++t)
            ucol[t] = soff[hp.hidx[t]] + rank[hp.hidx[t] * hp.H + e / hp.H];
    const int S = soff[T];
    // subset columns grouped by destination order mu = m - q, with per-lambda weights
    const int lmx = hp.lmax, ngrp = 2 * lmx + 1;
    auto isqrt = [](int v) {
        int r = (int)std::sqrt((double)v);
        while (r * r > v) --r;
        while ((r + 1) * (r + 1) <= v) ++r;
        return r;
    };
    std::vector<int32_t> gstart(ngrp + 1, 0), gpos(std::max(S, 1)), mu_of(std::max(S, 1));
    for (int t = 0; t < T; ++t) {
        const int p = isqrt(t), q = t - p * p - p;
        for (int s = soff[t]; s < soff[t + 1]; ++s) {
            const int a1 = smode[s] + 1, la = isqrt(a1), ma = a1 - la * la - la;
            mu_of[s] = ma - q;
            if (std::abs(mu_of[s]) > lmx) {
                set_error("plan: subset column feeds no destination order");
                return TG_INVALID_INPUT;
            }
            gstart[mu_of[s] + lmx + 1]++;
        }
    }
    for (int g = 0; g < ngrp; ++g) gstart[g + 1] += gstart[g];
    {
        std::vector<int32_t> fillg(gstart.begin(), gstart.end() - 1);
        for (int s = 0; s < S; ++s) {
            gpos[s] = fillg[mu_of[s] + lmx]++;
        }
    }
    std::vector<double> gw((size_t)std::max(S, 1) * lmx * 2, 0.0);
    for (size_t e = 0; e < HH; ++e) {
        const int b1 = (int)(e % hp.H) + 1, lam = isqrt(b1), mu = b1 - lam * lam - lam;
        for (int32_t t = hp.starts[e]; t < hp.starts[e + 1]; ++t) {
            const int s = ucol[t], kind = (t - hp.starts[e]) >= hp.nsame[e];
            if (mu_of[s] != mu) {
                set_error("plan: term order mismatch");
                return TG_INVALID_INPUT;
            }
            gw[((size_t)gpos[s] * lmx + lam - 1) * 2 + kind] += hp.w[t];
        }
    }
    std::vector<int32_t> stiles((size_t)2 * kGemmCols * (T + 1), 0);
    for (int w = 0; w < 2; ++w)
        for (int cols = 1; cols <= kGemmCols; ++cols) {
            int32_t *st = stiles.data() + (size_t)(w * kGemmCols + cols - 1) * (T + 1);
            for (int t = 0; t < T; ++t) {
                const int ncol = 2 * cols * (soff[t + 1] - soff[t]);
                st[t + 1] = st[t] + (ncol + kTileWidths[w] - 1) / kTileWidths[w];
            }
        }
    auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
    size_t off_starts = 0;
    size_t off_nsame = al(off_starts + sizeof(int32_t) * (HH + 1));
    size_t off_w = al(off_nsame + HH);
    size_t off_hidx = al(off_w + sizeof(double) * nterms);
    size_t off_refl = al(off_hidx + sizeof(uint16_t) * nterms);
    size_t off_ts = al(off_refl + sizeof(double) * hp.table_len);
    size_t off_te = al(off_ts + sizeof(int32_t) * (hp.table_len + 1));
    size_t off_tw = al(off_te + sizeof(int32_t) * nterms);
    size_t off_soff = al(off_tw + sizeof(double) * nterms);
    size_t off_smode = al(off_soff + sizeof(int32_t) * (T + 1));
    size_t off_ucol = al(off_smode + sizeof(int32_t) * std::max(S, 1));
    size_t off_stiles = al(off_ucol + sizeof(int32_t) * nterms);
    size_t off_gst = al(off_stiles + sizeof(int32_t) * stiles.size());
    size_t off_gpos = al(off_gst + sizeof(int32_t) * gstart.size());
    size_t off_gw = al(off_gpos + sizeof(int32_t) * gpos.size());
    size_t total = al(off_gw + sizeof(double) * gw.size());
    TG_TRY(buf.ensure(total));
    char *base = buf.as<char>();
    std::vector<char> host(total, 0);
    memcpy(host.data() + off_starts, hp.starts.data(), sizeof(int32_t) * (HH + 1));
    memcpy(host.data() + off_nsame, hp.nsame.data(), HH);
    memcpy(host.data() + off_w, hp.w.data(), sizeof(double) * nterms);
    memcpy(host.data() + off_hidx, hp.hidx.data(), sizeof(uint16_t) * nterms);
    memcpy(host.data() + off_refl, hp.refl.data(), sizeof(double) * hp.table_len);
    memcpy(host.data() + off_ts, tstarts.data(), sizeof(int32_t) * (hp.table_len + 1));
    memcpy(host.data() + off_te, tentry.data(), sizeof(int32_t) * nterms);
    memcpy(host.data() + off_tw, tw.data(), sizeof(double) * nterms);
    memcpy(host.data() + off_soff, soff.data(), sizeof(int32_t) * (T + 1));
    if (S) memcpy(host.data() + off_smode, smode.data(), sizeof(int32_t) * S);
    memcpy(host.data() + off_ucol, ucol.data(), sizeof(int32_t) * nterms);
    memcpy(host.data() + off_stiles, stiles.data(), sizeof(int32_t) * stiles.size());
    memcpy(host.data() + off_gst, gstart.data(), sizeof(int32_t) * gstart.size());
    memcpy(host.data() + off_gpos, gpos.data(), sizeof(int32_t) * gpos.size());
    memcpy(host.data() + off_gw, gw.data(), sizeof(double) * gw.size());
    TG_HIP(hipMemcpy(base, host.data(), total, hipMemcpyHostToDevice));
    dp.H = hp.H;
    dp.L2 = hp.L2;
    dp.T = hp.table_len;
    dp.starts = (const int32_t *)(base + off_starts);
    dp.nsame = (const uint8_t *)(base + off_nsame);
    dp.w = (const double *)(base + off_w);
    dp.hidx = (const uint16_t *)(base + off_hidx);
    dp.refl = (const double *)(base + off_refl);
    dp.tstarts = (const int32_t *)(base + off_ts);
    dp.tentry = (const int32_t *)(base + off_te);
    dp.tw = (const double *)(base + off_tw);
    dp.S = S;
    dp.soff = (const int32_t *)(base + off_soff);
    dp.smode = (const int32_t *)(base + off_smode);
    dp.ucol = (const int32_t *)(base + off_ucol);
    dp.stiles = (const int32_t *)(base + off_stiles);
    dp.lmax = lmx;
    dp.gstart = (const int32_t *)(base + off_gst);
    dp.gpos = (const int32_t *)(base + off_gpos);
    dp.gw = (const double *)(base + off_gw);
    if (ntiles) {
        ntiles->assign((size_t)2 * kGemmCols, 0);
        for (int w = 0; w < 2; ++w)
            for (int cols = 1; cols <= kGemmCols; ++cols)
                (*ntiles)[(size_t)w * kGemmCols + cols - 1] = stiles[(size_t)(w * kGemmCols + cols - 1) * (T + 1) + T];
    }
    return TG_OK;
}

int upload_rect_plan(const HostPlan &hp, DevBuf &buf, RectPlan &rp, hipStream_t s) {
    const size_t E = (size_t)hp.H * hp.Hs, nterms = hp.w.size();
    auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
    const size_t o_st = 0, o_ns = al(sizeof(int32_t) * (E + 1)), o_w = al(o_ns + E);
    const size_t o_h = al(o_w + sizeof(double) * nterms), o_r = al(o_h + sizeof(uint16_t) * nterms);
    const size_t total = al(o_r + sizeof(double) * hp.table_len);
    TG_TRY(buf.ensure(total));
    std::vector<char> host(total, 0);
    memcpy(host.data() + o_st, hp.starts.data(), sizeof(int32_t) * (E + 1));
    memcpy(host.data() + o_ns, hp.nsame.data(), E);
    memcpy(host.data() + o_w, hp.w.data(), sizeof(double) * nterms);
    memcpy(host.data() + o_h, hp.hidx.data(), sizeof(uint16_t) * nterms);
    memcpy(host.data() + o_r, hp.refl.data(), sizeof(double) * hp.table_len);
    char *base = buf.as<char>();
    TG_HIP(hipMemcpyAsync(base, host.data(), total, hipMemcpyHostToDevice, s));
    TG_HIP(hipStreamSynchronize(s));
    rp.Hd = hp.H;
    rp.Hs = hp.Hs;
    rp.L2 = hp.L2;
    rp.T = hp.table_len;
    rp.starts = (const int32_t *)(base + o_st);
    rp.nsame = (const uint8_t *)(base + o_ns);
    rp.w = (const double *)(base + o_w);
    rp.hidx = (const uint16_t *)(base + o_h);
    rp.refl = (const double *)(base + o_r);
    return TG_OK;
}

// y = (I - T C) x or (I - C^H T^H) x, split layout.
int apply_operator(tg_cluster *c, const cd *x, cd *y, int cols, int adjoint, int op) {
    const int64_t count = (int64_t)cols * c->n;
    TG_TRY(c->ws[6].ensure(sizeof(cd) * count * 2));
    cd *t1 = c->ws[6].as<cd>();
    cd *t2 = t1 + count;
    auto coupling = [&](const cd *in, cd *out) -> int {
        if (op == TG_OP_STORED) return coupling_stored(c, in, out, cols);
        if (op == TG_OP_MATRIX_FREE_C64 || op == TG_OP_MATRIX_FREE_C64_HI)
            return coupling_matrix_free32(c, in, out, cols, op == TG_OP_MATRIX_FREE_C64_HI);
        return coupling_matrix_free(c, in, out, cols, 0);
    };
    if (!adjoint) {
        TG_TRY(coupling(x, t1));
        TG_TRY(local_apply(c, t1, y, cols, 0, x));
        return TG_OK;
    }
    // C^H z = P conj(C (P conj(z))) with z = T^H x
    TG_TRY(local_apply(c, x, t1, cols, 1, nullptr));
    TG_TRY(reflect_conj(c, t1, t2, cols));
    TG_TRY(coupling(t2, t1));
    TG_TRY(reflect_conj(c, t1, t2, cols));
    TG_TRY(vec_axpby(c->stream, cd(1.0), x, cd(-1.0), t2, count));
    TG_HIP(hipMemcpyAsync(y, t2, sizeof(cd) * count, hipMemcpyDeviceToDevice, c->stream));
    return TG_OK;
}

static int check_cluster(const tg_cluster *c) {
    if (!c) {
        set_error("null cluster handle");
        return TG_INVALID_INPUT;
    }
    return TG_OK;
}

}  // namespace tg

using namespace tg;

// The translation plan depends on lmax alone (0.2 ms to build at lmax 3, 3.7 ms at
// lmax 6): built once per process and copied into each cluster.
static int cached_plan(int lmax, HostPlan &out) {
    static std::mutex mu;
    static std::map<int, HostPlan> cache;
    std::lock_guard<std::mutex> lock(mu);
    auto it = cache.find(lmax);
    if (it == cache.end()) {
        HostPlan hp;
        if (build_plan(lmax, hp)) return 1;
        it = cache.emplace(lmax, std::move(hp)).first;
    }
    out = it->second;
    return 0;
}

// The device plan (with its transposed plan, GEMM subsets and tile counts) depends on
// lmax alone too: uploaded once per (device, lmax) and shared by every cluster on
// that device, never freed (a few MB at lmax 12).
static int shared_device_plan(int lmax, const HostPlan &hp, DevPlan &out, std::vector<int> &ntiles) {
    struct Shared {
        DevPlan plan;
        std::vector<int> ntiles;
    };
    static std::mutex mu;
    static std::map<std::pair<int, int>, Shared> cache;
    int dev = 0;
    TG_HIP(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lock(mu);
    auto it = cache.find({dev, lmax});
    if (it == cache.end()) {
        DevBuf buf;
        Shared sh;
        TG_TRY(upload_plan(hp, buf, sh.plan, &sh.ntiles));
        buf.ptr = nullptr;  // owned by the plan from now on (never returned to the cache)
        buf.bytes = 0;
        it = cache.emplace(std::make_pair(dev, lmax), std::move(sh)).first;
    }
    out = it->second.plan;
    ntiles = it->second.ntiles;
    return TG_OK;
}

extern "C" {

const char *tg_last_error(void) { return g_error.c_str(); }

int tg_modes_per_sphere(int lmax) { return 2 * lmax * (lmax + 2); }

double tg_wigner3j(int j1, int j2, int j3, int m1, int m2, int m3) { return wigner3j(j1, j2, j3, m1, m2, m3); }

int tg_rotation_matrix(int lmax, double phi, double theta, double psi, tg_c128 *out) {
    if (lmax < 1 || lmax > kMaxLmax || !out || !std::isfinite(phi) || !std::isfinite(theta) || !std::isfinite(psi)) {
        set_error("require 1 <= lmax <= %d, finite Euler angles and an output", kMaxLmax);
        return TG_INVALID_INPUT;
    }
    const int M = 2 * lmax * (lmax + 2);
    std::fill(out, out + (size_t)M * M, tg_c128{0.0, 0.0});
    for (int l = 1; l <= lmax; ++l)
        for (int mp = -l; mp <= l; ++mp)
            for (int m = -l; m <= l; ++m) {
                // D^l_(m' m) = e^(-i m' phi) d^l_(m' m)(theta) e^(-i m psi)
                const double d = wigner_small_d(l, mp, m, theta), arg = -(mp * phi + m * psi);
                const tg_c128 v{d * std::cos(arg), d * std::sin(arg)};
                for (int pol = 0; pol < 2; ++pol) {
                    const int r = 2 * (l * l + l + mp - 1) + (1 - pol), c = 2 * (l * l + l + m - 1) + (1 - pol);
                    out[(size_t)r * M + c] = v;
                }
            }
    return TG_OK;
}

int tg_plan_terms(int lmax, int32_t *starts, uint8_t *nsame, double *w, uint16_t *hidx, int64_t *count) {
    HostPlan hp;
    if (build_plan(lmax, hp)) {
        set_error("require 1 <= lmax <= %d", kMaxLmax);
        return TG_INVALID_INPUT;
    }
    if (count) *count = (int64_t)hp.w.size();
    if (starts) memcpy(starts, hp.starts.data(), sizeof(int32_t) * hp.starts.size());
    if (nsame) memcpy(nsame, hp.nsame.data(), hp.nsame.size());
    if (w) memcpy(w, hp.w.data(), sizeof(double) * hp.w.size());
    if (hidx) memcpy(hidx, hp.hidx.data(), sizeof(uint16_t) * hp.hidx.size());
    return TG_OK;
}

int tg_cluster_create(int lmax, double k0, int n_spheres, const double *radii, const tg_c128 *epsilon,
                      const double *positions, void *stream, tg_cluster **out) {
    if (!out || !radii || !epsilon || !positions) {
        set_error("null argument");
        return TG_INVALID_INPUT;
    }
    *out = nullptr;
    if (lmax < 1 || lmax > kMaxLmax) {
        set_error("require 1 <= lmax <= %d", kMaxLmax);
        return TG_INVALID_INPUT;
    }
    // cluster/spheres.rs:99-133
    bool ok = std::isfinite(k0) && k0 > 0.0 && n_spheres > 0;
    for (int i = 0; ok && i < n_spheres; ++i) {
        ok = std::isfinite(radii[i]) && radii[i] > 0.0 && std::isfinite(epsilon[i].re) &&
             std::isfinite(epsilon[i].im);
        for (int a = 0; a < 3; ++a) ok = ok && std::isfinite(positions[3 * i + a]);
    }
    if (!ok) {
        set_error("require positive finite k0 and radii, and matching finite permittivities and positions");
        return TG_INVALID_INPUT;
    }
    double dmin = INFINITY;  // the closest pair of centres
    for (int i = 0; i < n_spheres; ++i)
        for (int j = 0; j < i; ++j) {
            double d = 0.0;
            for (int a = 0; a < 3; ++a) d = std::hypot(d, positions[3 * i + a] - positions[3 * j + a]);
            if (!(d > radii[i] + radii[j]) || !std::isfinite(d)) {
                set_error("spherical particles must have finite separation and must not touch or overlap");
                return TG_INVALID_INPUT;
            }
            dmin = std::min(dmin, d);
        }
    tg_cluster *c = new tg_cluster();
    cluster_born();
    c->dmin = dmin;
    c->lmax = lmax;
    c->N = n_spheres;
    c->M = 2 * lmax * (lmax + 2);
    c->H = lmax * (lmax + 2);
    c->n = (int64_t)c->N * c->M;
    c->nh = (int64_t)c->N * c->H;
    c->k0 = k0;
    c->stream = (hipStream_t)stream;
    c->radii.assign(radii, radii + n_spheres);
    c->positions.assign(positions, positions + 3 * n_spheres);
    for (int i = 0; i < n_spheres; ++i) c->eps.push_back(cd(epsilon[i].re, epsilon[i].im));
    int st = TG_OK;
    do {
        if (cached_plan(lmax, c->hplan)) {
            set_error("plan construction failed");
            st = TG_INVALID_INPUT;
            break;
        }
        if ((st = shared_device_plan(lmax, c->hplan, c->plan, c->gemm_ntiles))) break;
        std::vector<double> geo(6 * (size_t)n_spheres);
        memcpy(geo.data(), positions, sizeof(double) * 3 * n_spheres);
        memcpy(geo.data() + 3 * n_spheres, radii, sizeof(double) * n_spheres);
        for (int i = 0; i < n_spheres; ++i) {
            geo[4 * (size_t)n_spheres + 2 * i] = epsilon[i].re;
            geo[4 * (size_t)n_spheres + 2 * i + 1] = epsilon[i].im;
        }
        if ((st = c->b_pos.ensure(sizeof(double) * geo.size()))) break;
        // (geo outlives the stream synchronisation below)
        hipError_t e = hipMemcpyAsync(c->b_pos.ptr, geo.data(), sizeof(double) * geo.size(), hipMemcpyHostToDevice,
                                      c->stream);
        if (e != hipSuccess) {
            set_error("upload: %s", hipGetErrorString(e));
            st = TG_DEVICE_ERROR;
            break;
        }
        if ((st = c->b_mie.ensure(sizeof(cd) * 12 * (size_t)n_spheres * lmax))) break;
        if ((st = launch_mie(c))) break;
        std::vector<cd> mie(4 * (size_t)n_spheres * lmax);
        e = hipMemcpyAsync(mie.data(), c->b_mie.ptr, sizeof(cd) * mie.size(), hipMemcpyDeviceToHost, c->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
        if (e != hipSuccess) {
            set_error("Mie blocks: %s", hipGetErrorString(e));
            st = TG_DEVICE_ERROR;
            break;
        }
        for (auto &v : mie)
            if (!std::isfinite(v.re) || !std::isfinite(v.im)) {
                set_error("singular Mie coefficients");
                st = TG_SINGULAR;
                break;
            }
    } while (0);
    if (st) {
        tg_cluster_destroy(c);
        return st;
    }
    *out = c;
    return TG_OK;
}

int tg_cluster_destroy(tg_cluster *c) {
    if (!c) return TG_OK;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    (void)side_sync(c);
    // both streams are idle: the cluster's blocks go back to the cache clean
    CleanReleases clean;
    if (c->side) {
        (void)hipStreamDestroy(c->side);
        (void)hipEventDestroy(c->side_start);
        (void)hipEventDestroy(c->side_split);
        (void)hipEventDestroy(c->side_done);
        c->side = nullptr;
    }
    destroy_preconditioner(c);
    c->b_plan.release();
    c->b_pos.release();
    c->b_mie.release();
    c->b_C.release();
    c->b_tab.release();
    c->have_tables = false;
    c->b_tab32.release();
    c->have_tables32 = false;
    c->b_tabh.release();
    c->b_texp.release();
    c->have_tablesh = false;
    c->b_tabR.release();
    c->have_tablesR = false;
    c->b_tloc.release();
    learned_release(c);
    for (auto &w : c->ws) w.release();
    c->outer.release();
    c->pin.release();
    c->timer.collect();
    c->timer.destroy_events();
    delete c;
    cluster_died();
    return TG_OK;
}

int64_t tg_cluster_dimension(const tg_cluster *c) { return c ? c->n : 0; }

int tg_cluster_mie_blocks(tg_cluster *c, tg_c128 *out) {
    TG_TRY(check_cluster(c));
    TG_TRY(require_spheres(c, "the Mie blocks"));
    TG_HIP(hipMemcpyAsync(out, c->b_mie.ptr, sizeof(cd) * 4 * c->N * c->lmax, hipMemcpyDeviceToHost, c->stream));
    TG_HIP(hipStreamSynchronize(c->stream));
    return TG_OK;
}

int tg_pair_matrix(tg_cluster *c, int regular, tg_c128 *out) {
    TG_TRY(check_cluster(c));
    TG_TRY(launch_assemble_ref(c, regular, (cd *)out));
    TG_HIP(hipStreamSynchronize(c->stream));
    return TG_OK;
}

int tg_translation_blocks(int lmax, double k, int64_t count, const double *disp, int regular, tg_c128 *out,
                          void *stream) {
    HostPlan hp;
    if (build_plan(lmax, hp) || !(std::isfinite(k)) || count < 0) {
        set_error("invalid lmax, wavenumber or count");
        return TG_INVALID_INPUT;
    }
    DevBuf buf;
    DevPlan dp;
    int st = upload_plan(hp, buf, dp);
    if (!st) st = launch_translation_blocks(lmax, dp, k, count, disp, regular, (cd *)out, (hipStream_t)stream);
    if (!st) {
        hipError_t e = hipStreamSynchronize((hipStream_t)stream);
        if (e != hipSuccess) {
            set_error("%s", hipGetErrorString(e));
            st = TG_DEVICE_ERROR;
        }
    }
    buf.release();
    return st;
}

int tg_expansion_field(int lmax, double k0, const double *center, const tg_c128 *x, int cols, int64_t npoints,
                       const double *points, int regular, tg_c128 *out, void *stream) {
    HostPlan hp;
    if (lmax < 1 || lmax > kMaxLmax || !center || !x || !out || cols < 1 || npoints < 0 ||
        (npoints > 0 && !points) || !std::isfinite(k0) || !(k0 > 0) || !std::isfinite(center[0]) ||
        !std::isfinite(center[1]) || !std::isfinite(center[2]) || cached_plan(lmax, hp)) {
        set_error("invalid lmax (1 .. %d), wavenumber, centre, columns, points or buffers", kMaxLmax);
        return TG_INVALID_INPUT;
    }
    if (npoints == 0) return TG_OK;
    DevPlan dp;
    std::vector<int> ntiles;
    TG_TRY(shared_device_plan(lmax, hp, dp, ntiles));
    hipStream_t s = (hipStream_t)stream;
    DevBuf c;
    TG_TRY(c.ensure(sizeof(double) * 3));
    TG_HIP(hipMemcpyAsync(c.ptr, center, sizeof(double) * 3, hipMemcpyHostToDevice, s));
    const int64_t M = 2 * (int64_t)lmax * (lmax + 2);
    const int st = launch_field(1, lmax, k0, c.as<double>(), dp.refl, (const cd *)x, M, cols, npoints, points,
                                regular != 0, (cd *)out, s);
    if (st == TG_OK) {
        CleanReleases clean;  // launch_field synchronised the stream
        c.release();
    } else {
        (void)hipStreamSynchronize(s);
        c.release();
    }
    return st;
}

int tg_translation_blocks_tangent(int lmax, double k, int64_t count, const double *disp,
                                  const double *tangent_disp, double tangent_k, int regular, tg_c128 *out,
                                  void *stream) {
    HostPlan hp;
    if (lmax < 1 || lmax > kMaxLmax || cached_plan(lmax, hp) || !std::isfinite(k) || !(k > 0) ||
        !std::isfinite(tangent_k) || count < 0 || (count > 0 && (!disp || !tangent_disp || !out))) {
        set_error("invalid lmax, wavenumber, wavenumber tangent, count or buffers");
        return TG_INVALID_INPUT;
    }
    DevBuf buf, tables;
    DevPlan dp;
    hipStream_t s = (hipStream_t)stream;
    int st = upload_plan(hp, buf, dp);
    if (!st) st = tables.ensure(sizeof(cd) * (size_t)std::max<int64_t>(count, 1) * hp.table_len);
    if (!st)
        st = launch_translation_tangent(lmax, dp, k, count, disp, tangent_disp, tangent_k, regular, tables.as<cd>(),
                                        (cd *)out, s);
    if (!st && hipStreamSynchronize(s) != hipSuccess) {
        set_error("device failure in the translation tangents");
        st = TG_DEVICE_ERROR;
    }
    tables.release();
    buf.release();
    return st;
}

}  // extern "C"

// ----------------------------------------------------- allocation-free blocks --
// tg_translation_blocks_noalloc keeps its plan in `scratch`, a fixed device array of
// every GPU the library runs on, instead of a device allocation: the compact plan
// (starts, nsame, weights, harmonic indices, reflection factors) of lmax <= 12 needs
// at most 2.95 MiB.  In host code `scratch` names a host-side stand-in, never device
// memory: its device address comes from hipGetSymbolAddress, once per device, and
// only that address is used.  Every call on a GPU overwrites that GPU's copy, so
// calls that overlap on one GPU race.
static constexpr size_t kScratchBytes = (size_t)4 << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];
namespace {
constexpr int kScratchDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[kScratchDevices] = {};

// Device address of the current device's `scratch`, looked up on first use.
int scratch_address(void **out) {
    int dev = 0;
    TG_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kScratchDevices) {
        set_error("device ordinal %d beyond the %d devices the scratch table holds", dev, kScratchDevices);
        return TG_INVALID_INPUT;
    }
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            set_error("hipGetSymbolAddress(scratch) on device %d: %s", dev, hipGetErrorString(e));
            return TG_DEVICE_ERROR;
        }
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return TG_OK;
}
}  // namespace

extern "C" {

int tg_translation_blocks_noalloc(int lmax, double k, int64_t count, const double *disp, int regular, tg_c128 *out,
                                  void *stream) {
    HostPlan hp;
    if (lmax < 1 || lmax > kMaxLmax || cached_plan(lmax, hp) || !std::isfinite(k) || count < 0 ||
        (count > 0 && (!disp || !out))) {
        set_error("invalid lmax (1..%d), wavenumber, count or buffers", kMaxLmax);
        return TG_INVALID_INPUT;
    }
    const size_t HH = (size_t)hp.H * hp.H, nterms = hp.w.size();
    auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
    const size_t o_st = 0, o_ns = al(sizeof(int32_t) * (HH + 1)), o_w = al(o_ns + HH);
    const size_t o_h = al(o_w + sizeof(double) * nterms), o_r = al(o_h + sizeof(uint16_t) * nterms);
    const size_t total = al(o_r + sizeof(double) * hp.table_len);
    if (total > kScratchBytes) {
        set_error("the plan of lmax %d needs %zu bytes of scratch; the allocation-free path holds %zu", lmax, total,
                  kScratchBytes);
        return TG_INVALID_INPUT;
    }
    void *base = nullptr;
    TG_TRY(scratch_address(&base));
    std::vector<char> host(total, 0);
    memcpy(host.data() + o_st, hp.starts.data(), sizeof(int32_t) * (HH + 1));
    memcpy(host.data() + o_ns, hp.nsame.data(), HH);
    memcpy(host.data() + o_w, hp.w.data(), sizeof(double) * nterms);
    memcpy(host.data() + o_h, hp.hidx.data(), sizeof(uint16_t) * nterms);
    memcpy(host.data() + o_r, hp.refl.data(), sizeof(double) * hp.table_len);
    hipStream_t s = (hipStream_t)stream;
    char *b = (char *)base;
    TG_HIP(hipMemcpyAsync(b, host.data(), total, hipMemcpyHostToDevice, s));
    DevPlan dp{};
    dp.H = hp.H;
    dp.L2 = hp.L2;
    dp.T = hp.table_len;
    dp.starts = (const int32_t *)(b + o_st);
    dp.nsame = (const uint8_t *)(b + o_ns);
    dp.w = (const double *)(b + o_w);
    dp.hidx = (const uint16_t *)(b + o_h);
    dp.refl = (const double *)(b + o_r);
    int st = launch_translation_blocks(lmax, dp, k, count, disp, regular, (cd *)out, s);
    // the staging vector must outlive the copy
    hipError_t e = hipStreamSynchronize(s);
    if (!st && e != hipSuccess) {
        set_error("%s", hipGetErrorString(e));
        st = TG_DEVICE_ERROR;
    }
    return st;
}

int tg_translation_blocks_rect(int lmax_dst, int lmax_src, double k, int64_t count, const double *disp, int regular,
                               tg_c128 *out, void *stream) {
    HostPlan hp;
    if (build_plan_rect(lmax_dst, lmax_src, hp) || !(std::isfinite(k)) || count < 0 || (count > 0 && (!disp || !out))) {
        set_error("invalid lmax pair, wavenumber, count or buffers");
        return TG_INVALID_INPUT;
    }
    DevBuf buf;
    RectPlan rp{};
    hipStream_t s = (hipStream_t)stream;
    int st = upload_rect_plan(hp, buf, rp, s);
    if (!st) st = launch_translation_rect(rp, k, count, disp, regular, (cd *)out, s);
    if (!st && hipStreamSynchronize(s) != hipSuccess) {
        set_error("device failure in the translation blocks");
        st = TG_DEVICE_ERROR;
    }
    buf.release();
    return st;
}

int tg_cluster_assemble(tg_cluster *c, int storage) {
    TG_TRY(check_cluster(c));
    if (storage != TG_STORE_C64 && storage != TG_STORE_C128) {
        set_error("storage must be TG_STORE_C64 or TG_STORE_C128");
        return TG_INVALID_INPUT;
    }
    size_t elem = storage == TG_STORE_C64 ? 8 : 16;
    size_t bytes = elem * 2 * (size_t)c->nh * (size_t)c->nh;
    c->storage = TG_STORE_NONE;
    TG_TRY(c->b_C.ensure(bytes));
    TG_TRY(launch_assemble_split(c, storage, c->b_C.ptr));
    TG_HIP(hipStreamSynchronize(c->stream));
    c->storage = storage;
    return TG_OK;
}

int tg_cluster_release(tg_cluster *c) {
    TG_TRY(check_cluster(c));
    TG_HIP(hipStreamSynchronize(c->stream));
    TG_TRY(side_sync(c));
    destroy_preconditioner(c);
    c->b_C.release();
    c->b_tab.release();
    c->have_tables = false;
    c->b_tab32.release();
    c->have_tables32 = false;
    c->b_tabh.release();
    c->b_texp.release();
    c->have_tablesh = false;
    c->b_tabR.release();
    c->have_tablesR = false;
    c->storage = TG_STORE_NONE;
    learned_release(c);
    for (auto &w : c->ws) w.release();
    c->outer.release();
    return TG_OK;
}

int tg_cluster_reset(tg_cluster *c) {
    TG_TRY(check_cluster(c));
    TG_HIP(hipStreamSynchronize(c->stream));
    TG_TRY(side_sync(c));
    destroy_preconditioner(c);
    learned_invalidate(c);
    c->storage = TG_STORE_NONE;
    c->have_tables = false;
    c->have_tables32 = false;
    c->have_tablesh = false;
    c->have_tablesR = false;
    return TG_OK;
}

int tg_cluster_set_inner_relaxation(tg_cluster *c, double threshold) {
    TG_TRY(check_cluster(c));
    if (!(threshold >= 0.0) || !std::isfinite(threshold)) {
        set_error("relaxation threshold must be finite and >= 0, got %g", threshold);
        return TG_INVALID_INPUT;
    }
    c->relax_threshold = threshold;
    return TG_OK;
}

int tg_cluster_set_inner_gemm(tg_cluster *c, int kind) {
    TG_TRY(check_cluster(c));
    if (kind != TG_INNER_FP32 && kind != TG_INNER_SPLIT16) {
        set_error("inner GEMM kind must be TG_INNER_FP32 or TG_INNER_SPLIT16, got %d", kind);
        return TG_INVALID_INPUT;
    }
    TG_HIP(hipStreamSynchronize(c->stream));
    TG_TRY(side_sync(c));
    c->inner_gemm = kind;
    c->b_tab32.release();
    c->have_tables32 = false;
    c->b_tabh.release();
    c->b_texp.release();
    c->have_tablesh = false;
    return TG_OK;
}

int tg_cluster_set_table_limit(tg_cluster *c, int64_t bytes) {
    TG_TRY(check_cluster(c));
    if (bytes < 0) {
        set_error("table limit must be >= 0 bytes");
        return TG_INVALID_INPUT;
    }
    TG_HIP(hipStreamSynchronize(c->stream));
    TG_TRY(side_sync(c));
    c->table_limit = bytes;
    c->b_tab.release();
    c->have_tables = false;
    c->b_tab32.release();
    c->have_tables32 = false;
    c->b_tabh.release();
    c->b_texp.release();
    c->have_tablesh = false;
    c->b_tabR.release();
    c->have_tablesR = false;
    return TG_OK;
}

int tg_cluster_set_group_size(tg_cluster *c, int spheres) {
    TG_TRY(check_cluster(c));
    if (spheres < 1 || spheres > 16) {
        set_error("group size must be 1..16 spheres");
        return TG_INVALID_INPUT;
    }
    c->group_size = spheres;
    if (c->pc_kind == TG_PC_BLOCK_JACOBI) destroy_preconditioner(c);
    return TG_OK;
}

// reference-layout in/out wrappers around split-layout work
static int with_split(tg_cluster *c, const tg_c128 *x, tg_c128 *y, int cols,
                      const std::function<int(const cd *, cd *)> &body) {
    if (cols < 1) {
        set_error("cols must be positive");
        return TG_INVALID_INPUT;
    }
    const int64_t count = (int64_t)cols * c->n;
    TG_TRY(c->ws[7].ensure(sizeof(cd) * count * 2));
    cd *xs = c->ws[7].as<cd>(), *ys = xs + count;
    TG_TRY(ref_to_split(c, (const cd *)x, xs, cols));
    TG_TRY(body(xs, ys));
    TG_TRY(split_to_ref(c, ys, (cd *)y, cols));
    TG_HIP(hipStreamSynchronize(c->stream));
    return TG_OK;
}

int tg_cluster_coupling(tg_cluster *c, const tg_c128 *x, tg_c128 *y, int cols, int adjoint, int op) {
    TG_TRY(check_cluster(c));
    return with_split(c, x, y, cols, [&](const cd *xs, cd *ys) -> int {
        auto cp = [&](const cd *in, cd *out) -> int {
            if (op == TG_OP_STORED) return coupling_stored(c, in, out, cols);
            if (op == TG_OP_MATRIX_FREE_C64 || op == TG_OP_MATRIX_FREE_C64_HI)
                return coupling_matrix_free32(c, in, out, cols, op == TG_OP_MATRIX_FREE_C64_HI);
            return coupling_matrix_free(c, in, out, cols, 0);
        };
        if (!adjoint) return cp(xs, ys);
        const int64_t count = (int64_t)cols * c->n;
        TG_TRY(c->ws[6].ensure(sizeof(cd) * count * 2));
        cd *t1 = c->ws[6].as<cd>(), *t2 = t1 + count;
        TG_TRY(reflect_conj(c, xs, t1, cols));
        TG_TRY(cp(t1, t2));
        return reflect_conj(c, t2, ys, cols);
    });
}

int tg_cluster_overlap(tg_cluster *c, const tg_c128 *x, tg_c128 *y, int cols, int adjoint) {
    TG_TRY(check_cluster(c));
    return with_split(c, x, y, cols, [&](const cd *xs, cd *ys) -> int {
        if (!adjoint) return coupling_matrix_free(c, xs, ys, cols, 1);
        const int64_t count = (int64_t)cols * c->n;
        TG_TRY(c->ws[6].ensure(sizeof(cd) * count * 2));
        cd *t1 = c->ws[6].as<cd>(), *t2 = t1 + count;
        TG_TRY(reflect_conj(c, xs, t1, cols));
        TG_TRY(coupling_matrix_free(c, t1, t2, cols, 1));
        return reflect_conj(c, t2, ys, cols);
    });
}

int tg_cluster_local(tg_cluster *c, const tg_c128 *x, tg_c128 *y, int cols, int adjoint) {
    TG_TRY(check_cluster(c));
    return with_split(c, x, y, cols,
                      [&](const cd *xs, cd *ys) -> int { return local_apply(c, xs, ys, cols, adjoint, nullptr); });
}

int tg_cluster_apply(tg_cluster *c, const tg_c128 *x, tg_c128 *y, int cols, int adjoint, int op) {
    TG_TRY(check_cluster(c));
    return with_split(c, x, y, cols,
                      [&](const cd *xs, cd *ys) -> int { return apply_operator(c, xs, ys, cols, adjoint, op); });
}

int tg_cluster_plane_wave(tg_cluster *c, const double *dir, int pol, tg_c128 *out) {
    TG_TRY(check_cluster(c));
    if (!dir || !(std::isfinite(dir[0]) && std::isfinite(dir[1]) && std::isfinite(dir[2])) ||
        (dir[0] == 0 && dir[1] == 0 && dir[2] == 0) || (pol != 0 && pol != 1)) {
        set_error("plane wave needs a finite nonzero direction and pol 0 or 1");
        return TG_INVALID_INPUT;
    }
    TG_TRY(c->ws[7].ensure(sizeof(cd) * c->n));
    cd *tmp = c->ws[7].as<cd>();
    TG_TRY(plane_wave(c, dir, pol, tmp));
    TG_TRY(split_to_ref(c, tmp, (cd *)out, 1));
    TG_HIP(hipStreamSynchronize(c->stream));
    return TG_OK;
}

}  // extern "C"

namespace tg {

// Solve in the split layout: b (already T-applied for forward solves) -> x.
int solve_split(tg_cluster *c, const cd *b, cd *x, int cols, const tg_gmres_options &o, int adjoint,
                std::vector<tg_convergence> &conv) {
    const int64_t count = (int64_t)cols * c->n;
    GmresParams prm;
    prm.rtol = o.rtol;
    prm.atol = o.atol;
    prm.restart = o.restart;
    prm.max_iterations = o.max_iterations;
    OpFn pc_fn;
    const OpFn *pc = nullptr;
    // the inner coupling's tables build on the side stream meanwhile
    if (o.mixed_precision == 2 && o.preconditioner != TG_PC_NONE) TG_TRY(prefetch_inner_tables(c));
    if (o.preconditioner != TG_PC_NONE) {
        TG_TRY(ensure_preconditioner(c, o.preconditioner, o.mixed_precision ? TG_STORE_C64 : TG_STORE_C128));
        pc_fn = [c, adjoint](const cd *in, cd *out, int nc) -> int { return apply_preconditioner(c, in, out, nc, adjoint); };
        pc = &pc_fn;
    }
    if (!o.mixed_precision) {
        const int op = c->storage == TG_STORE_C128 ? TG_OP_STORED : TG_OP_MATRIX_FREE_F64;
        OpFn A = [c, adjoint, op](const cd *in, cd *out, int nc) -> int {
            return apply_operator(c, in, out, nc, adjoint, op);
        };
        int its = 0;
        int st = gmres_run(c, A, pc, b, x, cols, prm, conv, &its);
        return st;
    }
    // Mixed precision: inner GMRES on a c64 operator (the stored c64 coupling, or with
    // mixed_precision == 2 the matrix-free coupling on FP32 MFMA), outer FP64 residuals.
    const int op32 = o.mixed_precision == 2 ? TG_OP_MATRIX_FREE_C64 : TG_OP_STORED;
    if (op32 == TG_OP_STORED && c->storage != TG_STORE_C64) TG_TRY(tg_cluster_assemble(c, TG_STORE_C64));
    OpFn A32 = [c, adjoint, op32](const cd *in, cd *out, int nc) -> int {
        return apply_operator(c, in, out, nc, adjoint, op32);
    };
    // kept with the cluster: a fresh block from the buffer cache would synchronise the
    // device, and with it the side stream's table build
    DevBuf &outer = c->outer;
    TG_TRY(outer.ensure(sizeof(cd) * count * 3));
    cd *r = outer.as<cd>(), *d = r + count, *ax = d + count;
    std::vector<double> bn(cols), rn(cols);
    TG_TRY(vec_norms(c, b, c->n, cols, c->n, bn.data()));
    TG_HIP(hipMemsetAsync(x, 0, sizeof(cd) * count, c->stream));
    TG_HIP(hipMemcpyAsync(r, b, sizeof(cd) * count, hipMemcpyDeviceToDevice, c->stream));
    conv.assign(cols, tg_convergence{0, 0, 0.0, 0.0});
    int iterations = 0, refinements = 0;
    int st = TG_OK;
    while (true) {
        TG_TRY(vec_norms(c, r, c->n, cols, c->n, rn.data()));
        double worst_ratio = 0.0;
        bool done = true;
        for (int k = 0; k < cols; ++k) {
            double tol = std::max(o.atol, o.rtol * bn[k]);
            conv[k].rhs_norm = bn[k];
            conv[k].residual_norm = rn[k];
            conv[k].iterations = iterations;
            conv[k].refinements = refinements;
            if (!std::isfinite(rn[k])) {
                set_error("non-finite residual in mixed-precision refinement");
                return TG_NOT_CONVERGED;
            }
            if (rn[k] > tol) done = false;
            if (rn[k] > 0) worst_ratio = std::max(worst_ratio, rn[k] > 0 ? tol / rn[k] : 1.0);
        }
        if (done) break;
        if (iterations >= o.max_iterations || refinements >= 20) {
            set_error("mixed-precision GMRES did not converge after %d iterations (%d refinements)", iterations,
                      refinements);
            return TG_NOT_CONVERGED;
        }
        // inner relative tolerance: half of what the outer still needs, but not below
        // 1e-6 of the current residual: the c64 operator differs from the FP64 one at
        // ~1e-7 relative, so one more FP64 residual is needed in any case
        double eta = 1e-6;
        double need = 0.5 * worst_ratio;
        for (int k = 0; k < cols; ++k) {
            double tol = std::max(o.atol, o.rtol * bn[k]);
            if (rn[k] > tol) need = std::min(need, 0.5 * tol / rn[k]);
        }
        if (need > eta) eta = need;
        GmresParams ip = prm;
        ip.rtol = eta;
        ip.atol = 0.0;
        ip.inner = true;
        ip.max_iterations = o.max_iterations - iterations;
        std::vector<tg_convergence> ic;
        int its = 0;
        st = gmres_run(c, A32, pc, r, d, cols, ip, ic, &its);
        iterations += its;
        if (st != TG_OK && st != TG_NOT_CONVERGED) break;
        st = TG_OK;
        TG_TRY(vec_axpby(c->stream, cd(1.0), d, cd(1.0), x, count));
        ++refinements;
        const int op64 = TG_OP_MATRIX_FREE_F64;
        TG_TRY(apply_operator(c, x, ax, cols, adjoint, op64));
        TG_HIP(hipMemcpyAsync(r, b, sizeof(cd) * count, hipMemcpyDeviceToDevice, c->stream));
        TG_TRY(vec_axpby(c->stream, cd(-1.0), ax, cd(1.0), r, count));
        if (its == 0) {
            set_error("mixed-precision inner solve made no progress");
            st = TG_NOT_CONVERGED;
            break;
        }
    }
    return st;
}

}  // namespace tg

extern "C" {

int tg_cluster_solve(tg_cluster *c, const tg_c128 *incident, int cols, const tg_gmres_options *o, int adjoint,
                     tg_c128 *x, tg_convergence *conv_host) {
    TG_TRY(check_cluster(c));
    if (!o || !incident || !x || cols < 1) {
        set_error("null argument or no columns");
        return TG_INVALID_INPUT;
    }
    // linalg/gmres.rs:64-78
    if (!std::isfinite(o->rtol) || !std::isfinite(o->atol) || o->rtol < 0 || o->atol < 0 || o->rtol + o->atol == 0 ||
        o->restart <= 0 || o->max_iterations <= 0) {
        set_error("GMRES needs finite nonnegative tolerances, a positive tolerance, restart and iteration limit");
        return TG_INVALID_INPUT;
    }
    const int64_t count = (int64_t)cols * c->n;
    DevBuf io;
    TG_TRY(io.ensure(sizeof(cd) * count * 3));
    cd *a = io.as<cd>(), *b = a + count, *xs = b + count;
    int st = ref_to_split(c, (const cd *)incident, a, cols);
    if (!st) st = adjoint ? (hipMemcpyAsync(b, a, sizeof(cd) * count, hipMemcpyDeviceToDevice, c->stream) == hipSuccess
                                 ? TG_OK
                                 : TG_DEVICE_ERROR)
                          : local_apply(c, a, b, cols, 0, nullptr);  // rhs = T a
    std::vector<tg_convergence> conv;
    if (!st) st = solve_split(c, b, xs, cols, *o, adjoint, conv);
    if (!st || st == TG_NOT_CONVERGED) {
        int s2 = split_to_ref(c, xs, (cd *)x, cols);
        if (!st) st = s2;
    }
    if (hipStreamSynchronize(c->stream) != hipSuccess && !st) st = TG_DEVICE_ERROR;
    if (conv_host)
        for (int k = 0; k < cols && k < (int)conv.size(); ++k) conv_host[k] = conv[k];
    io.release();
    return st;
}

int tg_cluster_cross_sections(tg_cluster *c, const tg_c128 *incident, const tg_c128 *x, double flux, double *sca,
                              double *ext) {
    TG_TRY(check_cluster(c));
    if (!std::isfinite(flux) || flux <= 0) {
        set_error("require finite incident coefficients and positive flux");
        return TG_INVALID_INPUT;
    }
    const int64_t n = c->n;
    DevBuf io;
    TG_TRY(io.ensure(sizeof(cd) * n * 3));
    cd *xs = io.as<cd>(), *rx = xs + n, *as = rx + n;
    int st = ref_to_split(c, (const cd *)x, xs, 1);
    if (!st) st = ref_to_split(c, (const cd *)incident, as, 1);
    if (!st) st = coupling_matrix_free(c, xs, rx, 1, 1);  // R x (self blocks identity)
    cd d1, d2;
    if (!st) st = vec_dots(c, xs, 0, 1, rx, n, 1, n, n, &d1);
    if (!st) st = vec_dots(c, as, 0, 1, xs, n, 1, n, n, &d2);
    io.release();
    if (st) return st;
    const double w = 1.0 / (c->k0 * c->k0);
    *sca = d1.re * w * 0.5 / flux;
    *ext = -d2.re * w * 0.5 / flux;
    return TG_OK;
}

int tg_cluster_timing_enable(tg_cluster *c, int enable) {
    TG_TRY(check_cluster(c));
    c->timer.enabled = enable != 0;
    return TG_OK;
}

int tg_cluster_timing_get(tg_cluster *c, const char *kernel, double *total_ms, int64_t *launches) {
    TG_TRY(check_cluster(c));
    TG_TRY(c->timer.collect());
    auto it = c->timer.totals.find(kernel);
    *total_ms = it == c->timer.totals.end() ? 0.0 : it->second.first;
    *launches = it == c->timer.totals.end() ? 0 : it->second.second;
    return TG_OK;
}

int tg_cluster_timing_reset(tg_cluster *c) {
    TG_TRY(check_cluster(c));
    TG_TRY(c->timer.collect());
    c->timer.totals.clear();
    return TG_OK;
}

}  // extern "C"
