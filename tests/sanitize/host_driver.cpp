// Host-side sanitizer driver for libhipsw (built with ASan+UBSan or TSan on the
// host pass only; the kernels are not instrumented).  Exercises the host code
// that runs concurrently or touches pinned staging: the one-shot entry points
// from several threads over the engine's stream-context pool, the autotuner's
// table written while other threads plan, the allocation-free entry point,
// injected failures, and the serving loop's worker threads with concurrent
// submitters.  Every result is compared with the same call run alone first.
// The chain driver's passes and the CIGAR route / finish steps run first, on
// the host alone, from several threads.
// Exit 0 when everything matched, 1 on a mismatch, 2 without a device (after
// the host-only steps matched).
#include <algorithm>
#include <atomic>
#include <cstdio>
#include <cstring>
#include <thread>
#include <vector>

#include "hipsw.h"

extern "C" {
int hipGetDeviceCount(int *count);            // HIP runtime (hipError_t is an int enum)
const char *hipGetErrorString(int err);
void *wl_generate(int kind, uint64_t seed, int64_t n);
void *wl_field(void *s, int which);
int64_t wl_count(void *s, int which);
void wl_free(void *s);
uint8_t *wl_reference(uint64_t seed, int64_t n, int copies, double div);
void wl_pack_pac(const uint8_t *ref, int64_t n, uint8_t *pac);
void wl_free_bytes(uint8_t *p);
void *wl_chunk(const uint8_t *ref, int64_t gl, uint64_t seed, int64_t pairs);
void *wl_chunk_field(void *c, int kind, int which);
int64_t wl_chunk_count(void *c, int kind, int which);
void wl_chunk_free(void *c);
}

namespace {

std::atomic<int> g_bad{0};

void expect(bool ok, const char *what)
{
    if (!ok) {
        std::fprintf(stderr, "MISMATCH: %s\n", what);
        g_bad++;
    }
}

struct Jobs {
    void *h;
    int n;
    const uint8_t *q, *t;
    const uint32_t *qo, *to;
    const uint16_t *ql, *tl;
    const int32_t *h0, *w, *eb, *ts;
};

Jobs gen(int kind, uint64_t seed, int n)
{
    Jobs j;
    j.h = wl_generate(kind, seed, n);
    j.n = n;
    j.q = (const uint8_t *)wl_field(j.h, 0);
    j.qo = (const uint32_t *)wl_field(j.h, 1);
    j.ql = (const uint16_t *)wl_field(j.h, 2);
    j.t = (const uint8_t *)wl_field(j.h, 3);
    j.to = (const uint32_t *)wl_field(j.h, 4);
    j.tl = (const uint16_t *)wl_field(j.h, 5);
    j.h0 = (const int32_t *)wl_field(j.h, 6);
    j.w = (const int32_t *)wl_field(j.h, 7);
    j.eb = (const int32_t *)wl_field(j.h, 8);
    j.ts = (const int32_t *)wl_field(j.h, 9);
    return j;
}

hipsw_scoring scoring()
{
    hipsw_scoring s;
    for (int k = 0; k < 5; ++k)
        for (int x = 0; x < 5; ++x) s.mat[k * 5 + x] = (k == 4 || x == 4) ? -1 : (k == x ? 1 : -4);
    s.o_ins = s.o_del = 6;
    s.e_ins = s.e_del = 1;
    s.zdrop = 100;
    return s;
}

int ext(hipsw_engine *e, const Jobs &j, int mode, std::vector<int32_t> &out, bool noalloc = false)
{
    const hipsw_scoring sc = scoring();
    out.assign((size_t)j.n * 6, 0);
    auto fn = noalloc ? hipsw_banded_extend_noalloc : hipsw_banded_extend;
    return fn(e, j.n, j.q, j.qo, j.ql, j.t, j.to, j.tl, j.h0, j.w, j.eb, &sc, mode, out.data(), nullptr);
}

int nw(hipsw_engine *e, const Jobs &j, std::vector<int32_t> &score, std::vector<uint32_t> &cig)
{
    const hipsw_scoring sc = scoring();
    score.assign(j.n, 0);
    cig.assign((size_t)j.n * 512, 0);
    std::vector<int32_t> ncig(j.n), flags(j.n);
    return hipsw_global_nw(e, j.n, j.q, j.qo, j.ql, j.t, j.to, j.tl, j.w, j.ts, 3, 512, &sc,
                           HIPSW_MODE_BATCHED, score.data(), cig.data(), ncig.data(), flags.data(),
                           nullptr, nullptr);
}

int loc(hipsw_engine *e, const Jobs &j, std::vector<int32_t> &out)
{
    const hipsw_scoring sc = scoring();
    out.assign((size_t)j.n * 6, 0);
    std::vector<int32_t> ok(j.n);
    return hipsw_local_sw(e, j.n, j.q, j.qo, j.ql, j.t, j.to, j.tl, 19, 1, 4, &sc, out.data(),
                          ok.data(), nullptr);
}

struct ChunkView {
    void *h;
    hipsw_chunk c;
};

ChunkView chunk(const uint8_t *ref, int64_t gl, uint64_t seed, int pairs)
{
    ChunkView v;
    v.h = wl_chunk(ref, gl, seed, pairs);
    hipsw_chunk &c = v.c;
    std::memset(&c, 0, sizeof c);
    auto f = [&](int kind, int which) { return wl_chunk_field(v.h, kind, which); };
    c.n_ext = (int32_t)wl_chunk_count(v.h, 0, 1);
    c.ext_q = (const uint8_t *)f(0, 0); c.ext_q_off = (const uint32_t *)f(0, 1);
    c.ext_q_len = (const uint16_t *)f(0, 2); c.ext_t_pos = (const int64_t *)f(0, 3);
    c.ext_t_len = (const uint16_t *)f(0, 4); c.ext_t_rev = (const uint8_t *)f(0, 5);
    c.ext_h0 = (const int32_t *)f(0, 6); c.ext_w = (const int32_t *)f(0, 7);
    c.ext_eb = (const int32_t *)f(0, 8);
    c.n_loc = (int32_t)wl_chunk_count(v.h, 1, 1);
    c.loc_q = (const uint8_t *)f(1, 0); c.loc_q_off = (const uint32_t *)f(1, 1);
    c.loc_q_len = (const uint16_t *)f(1, 2); c.loc_t_pos = (const int64_t *)f(1, 3);
    c.loc_t_len = (const uint16_t *)f(1, 4); c.loc_min_sub = 19;
    c.n_nw = (int32_t)wl_chunk_count(v.h, 2, 1);
    c.nw_q = (const uint8_t *)f(2, 0); c.nw_q_off = (const uint32_t *)f(2, 1);
    c.nw_q_len = (const uint16_t *)f(2, 2); c.nw_t_pos = (const int64_t *)f(2, 3);
    c.nw_t_len = (const uint16_t *)f(2, 4); c.nw_w = (const int32_t *)f(2, 6);
    c.nw_true_score = (const int32_t *)f(2, 7);
    return v;
}

struct ChunkOut {
    std::vector<int32_t> ext, loc, lok, nws, nwn, nwf;
    std::vector<uint32_t> cig;
    int status = 0;
};

int wait_chunk(hipsw_server *srv, int64_t t, const hipsw_chunk &c, ChunkOut &o)
{
    o.ext.assign((size_t)c.n_ext * 6, 0); o.loc.assign((size_t)c.n_loc * 6, 0);
    o.lok.assign(c.n_loc, 0); o.nws.assign(c.n_nw, 0); o.nwn.assign(c.n_nw, 0);
    o.nwf.assign(c.n_nw, 0); o.cig.assign((size_t)c.n_nw * 512, 0);
    hipsw_chunk_result r;
    std::memset(&r, 0, sizeof r);
    r.ext_out = o.ext.data(); r.loc_out = o.loc.data(); r.loc_ok = o.lok.data();
    r.nw_score = o.nws.data(); r.nw_cigar = o.cig.data(); r.nw_n_cigar = o.nwn.data();
    r.nw_flags = o.nwf.data();
    o.status = hipsw_server_wait(srv, t, &r);
    return o.status;
}

bool same(const ChunkOut &a, const ChunkOut &b)
{
    return a.ext == b.ext && a.loc == b.loc && a.lok == b.lok && a.nws == b.nws && a.nwn == b.nwn &&
           a.cig == b.cig;
}

// Host-only steps (no device): the chain driver's pass state machine and the
// CIGAR route / finish steps, driven with synthetic extension and NW results,
// each thread on its own objects; every thread's output must equal the
// single-thread output.
struct HostOut {
    std::vector<hipsw_region> regs;
    std::vector<int32_t> nreg, score, nm, ncig, flags;
    std::vector<int64_t> pos;
    std::vector<uint32_t> cig;
    std::vector<char> md;
    bool operator==(const HostOut &o) const
    {
        return nreg == o.nreg && score == o.score && nm == o.nm && ncig == o.ncig &&
               flags == o.flags && pos == o.pos && cig == o.cig && md == o.md &&
               std::memcmp(regs.data(), o.regs.data(), regs.size() * sizeof(hipsw_region)) == 0;
    }
};

HostOut host_steps(uint32_t seed)
{
    HostOut o;
    uint64_t x = seed * 0x9E3779B97F4A7C15ull + 1;
    auto rnd = [&](uint32_t m) { x ^= x << 13; x ^= x >> 7; x ^= x << 17; return (uint32_t)(x % m); };
    const int n_reads = 64, rl = 150, n_chains = 96;
    std::vector<uint8_t> reads((size_t)n_reads * rl);
    for (auto &b : reads) b = (uint8_t)rnd(4);
    std::vector<uint32_t> read_off(n_reads), chain_read(n_chains), off(n_chains + 1, 0);
    std::vector<int32_t> read_len(n_reads, rl);
    for (int r = 0; r < n_reads; ++r) read_off[r] = (uint32_t)(r * rl);
    std::vector<int64_t> rbeg;
    std::vector<int32_t> qbeg, len, score;
    for (int c = 0; c < n_chains; ++c) {
        chain_read[c] = (uint32_t)(c % n_reads);
        const int ns = 1 + (int)rnd(4);
        const int64_t base = 100000 + (int64_t)rnd(1000000);
        for (int k = 0; k < ns; ++k) {
            const int qb = (int)rnd(120), l = 19 + (int)rnd(12);
            rbeg.push_back(base + qb);
            qbeg.push_back(qb);
            len.push_back(std::min(l, rl - qb));
            score.push_back(std::min(l, rl - qb));
        }
        off[c + 1] = (uint32_t)rbeg.size();
    }
    const hipsw_scoring sc = scoring();
    const hipsw_ext_opts opt = {1, 4, 100, 5, 5, 19};
    hipsw_chain_driver *d = hipsw_chain_begin(n_chains, off.data(), rbeg.data(), qbeg.data(), len.data(),
                                              score.data(), chain_read.data(), n_reads, reads.data(),
                                              read_off.data(), read_len.data(), &sc, &opt);
    expect(d != nullptr, "chain begin");
    for (int p = 0; d && p < 4; ++p) {
        hipsw_chain_pass ps;
        expect(hipsw_chain_next(d, &ps) == HIPSW_OK && ps.index == p, "chain next");
        std::vector<int32_t> res((size_t)ps.n * 6);
        for (int k = 0; k < ps.n; ++k) {            // synthetic SWResult rows
            const int32_t ql = ps.q_len[k], tl = ps.t_len[k];
            res[k * 6 + 0] = ps.h0[k] + (k % 5 ? (int32_t)ql / 2 : 0);
            res[k * 6 + 1] = ql / 2;
            res[k * 6 + 2] = std::min(tl, ql / 2 + 1);
            res[k * 6 + 3] = std::min(tl, ql);
            res[k * 6 + 4] = k % 3 ? ps.h0[k] + ql - 3 : 0;
            res[k * 6 + 5] = (int32_t)(k % 90);
        }
        expect(hipsw_chain_feed(d, res.data()) == HIPSW_OK, "chain feed");
    }
    o.regs.assign(rbeg.size(), hipsw_region{});
    o.nreg.assign(n_chains, 0);
    expect(d && hipsw_chain_finish(d, o.regs.data(), o.nreg.data()) == HIPSW_OK, "chain finish");
    hipsw_chain_free(d);
    // CIGAR route and finish over synthetic regions, NW rows "qlenM" or "1I..M"
    const int n = 80, cap = 64, md_cap = 256;
    std::vector<uint8_t> q, r;
    std::vector<uint32_t> qo(n), ro(n);
    std::vector<uint16_t> ql(n), rlen(n);
    std::vector<int32_t> qb(n), qe(n), rdl(n, rl), ts(n), iw(n, 5);
    std::vector<uint8_t> rev(n);
    std::vector<int64_t> rp(n);
    for (int k = 0; k < n; ++k) {
        qo[k] = (uint32_t)q.size();
        ro[k] = (uint32_t)r.size();
        ql[k] = (uint16_t)(20 + rnd(100));
        rlen[k] = (uint16_t)(ql[k] + (k % 4 == 0 ? rnd(5) : 0));
        for (int i = 0; i < ql[k]; ++i) q.push_back((uint8_t)rnd(4));
        for (int i = 0; i < rlen[k]; ++i) r.push_back(i < ql[k] && rnd(10) ? q[qo[k] + i] : (uint8_t)rnd(4));
        qb[k] = (int)rnd(10);
        qe[k] = qb[k] + ql[k];
        rdl[k] = qe[k] + (int)rnd(10);
        ts[k] = ql[k] - (int)rnd(12);
        rev[k] = (uint8_t)(k & 1);
        rp[k] = 5000 + k * 300;
    }
    std::vector<int32_t> w(n);
    std::vector<uint8_t> route(n);
    expect(hipsw_cigar_route(n, ql.data(), rlen.data(), qb.data(), qe.data(), ts.data(), iw.data(),
                             &sc, 1, 4, w.data(), route.data()) == HIPSW_OK, "cigar route");
    std::vector<int32_t> ns, nn, nf;
    std::vector<uint32_t> ops, noff(1, 0);
    for (int k = 0; k < n; ++k) {
        if (route[k] != 1) continue;
        ns.push_back(ts[k]);
        if (rlen[k] > ql[k]) {
            ops.push_back((uint32_t)(rlen[k] - ql[k]) << 4 | 2u);
            ops.push_back((uint32_t)ql[k] << 4);
            nn.push_back(2);
        } else {
            ops.push_back((uint32_t)ql[k] << 4);
            nn.push_back(1);
        }
        nf.push_back(0);
        noff.push_back((uint32_t)ops.size());
    }
    if (ns.empty()) { ns.push_back(0); nn.push_back(0); nf.push_back(0); ops.push_back(0); }
    o.score.assign(n, 0); o.nm.assign(n, 0); o.ncig.assign(n, 0); o.flags.assign(n, 0);
    o.pos.assign(n, 0); o.cig.assign((size_t)n * cap, 0); o.md.assign((size_t)n * md_cap, 0);
    expect(hipsw_cigar_finish(n, q.data(), qo.data(), ql.data(), r.data(), ro.data(), rlen.data(),
                              qb.data(), qe.data(), rdl.data(), rev.data(), ts.data(), rp.data(),
                              route.data(), ns.data(), ops.data(), noff.data(), nn.data(), nf.data(),
                              512, cap, md_cap, o.score.data(), o.nm.data(), o.pos.data(),
                              o.cig.data(), o.ncig.data(), o.md.data(), o.flags.data()) == HIPSW_OK,
           "cigar finish");
    return o;
}

void host_only_steps()
{
    std::puts("step: host-only chain passes and CIGAR route / finish from 8 threads");
    std::fflush(stdout);
    std::vector<HostOut> ref;
    for (uint32_t s = 0; s < 4; ++s) ref.push_back(host_steps(s));
    std::vector<std::thread> th;
    for (int t = 0; t < 8; ++t)
        th.emplace_back([&, t] {
            for (int it = 0; it < 3; ++it) expect(host_steps((uint32_t)(t % 4)) == ref[t % 4], "host steps thread");
        });
    for (auto &x : th) x.join();
}

}  // namespace

int main()
{
    host_only_steps();
    hipsw_engine *eng = hipsw_engine_create(0);
    if (!eng) {
        int nd = -1;
        const int e = hipGetDeviceCount(&nd);
        std::printf("no HIP device (hipGetDeviceCount: %d devices, %s); host-only steps: %d mismatches\n",
                    nd, hipGetErrorString(e), g_bad.load());
        return g_bad ? 1 : 2;
    }
    Jobs E = gen(4, 71, 6000), U = gen(2, 72, 4000), N = gen(3, 73, 3000), L = gen(5, 74, 2000);
    std::vector<int32_t> rE, rEp, rU, rN, rL;
    std::vector<uint32_t> rNc;
    expect(ext(eng, E, HIPSW_MODE_BATCHED, rE) == HIPSW_OK, "extension alone");
    expect(ext(eng, E, HIPSW_MODE_PLAIN, rEp) == HIPSW_OK && rEp == rE, "plain == batched");
    expect(ext(eng, U, HIPSW_MODE_BATCHED, rU) == HIPSW_OK, "uniform alone");
    expect(nw(eng, N, rN, rNc) == HIPSW_OK, "nw alone");
    expect(loc(eng, L, rL) == HIPSW_OK, "local alone");
    std::puts("step: one-shot calls from 8 threads, autotuner table written meanwhile");
    std::fflush(stdout);
    {
        std::vector<std::thread> th;
        for (int t = 0; t < 8; ++t)
            th.emplace_back([&, t] {
                for (int it = 0; it < 3; ++it) {
                    std::vector<int32_t> o, s;
                    std::vector<uint32_t> c;
                    switch (t % 5) {
                    case 0: expect(ext(eng, E, HIPSW_MODE_BATCHED, o) == 0 && o == rE, "thread ext"); break;
                    case 1: expect(ext(eng, U, HIPSW_MODE_PLAIN, o) == 0 && o == rU, "thread ext plain"); break;
                    case 2: expect(nw(eng, N, s, c) == 0 && s == rN && c == rNc, "thread nw"); break;
                    case 3: expect(loc(eng, L, o) == 0 && o == rL, "thread local"); break;
                    default: expect(ext(eng, U, HIPSW_MODE_BATCHED, o, true) == 0 && o == rU, "thread noalloc");
                    }
                }
            });
        th.emplace_back([&] {                 // layout changes while others plan
            for (int it = 0; it < 50; ++it) {
                hipsw_ext_set_layout(eng, 1, it % 2 ? 1 : 0, it % 2 ? 5 : 0);
                hipsw_ext_set_layout16(eng, 2, it % 3 == 0 ? 6 : 0);
            }
            hipsw_ext_set_layout(eng, 1, 0, 0);
            hipsw_ext_set_layout16(eng, 2, 0);
        });
        for (auto &x : th) x.join();
    }
    std::puts("step: injected faults");
    std::fflush(stdout);
    for (int site = HIPSW_FAULT_ALLOC; site <= HIPSW_FAULT_COPY; ++site) {
        std::vector<int32_t> o, s;
        std::vector<uint32_t> c;
        hipsw_fault_inject(eng, site, 1);
        expect(ext(eng, E, HIPSW_MODE_BATCHED, o) == HIPSW_E_DEVICE, "ext fault status");
        hipsw_fault_inject(eng, site, 1);
        expect(nw(eng, N, s, c) == HIPSW_E_DEVICE, "nw fault status");
        hipsw_fault_inject(eng, site, 1);
        expect(loc(eng, L, o) == HIPSW_E_DEVICE, "local fault status");
        hipsw_fault_inject(eng, HIPSW_FAULT_NONE, 0);
        expect(ext(eng, E, HIPSW_MODE_BATCHED, o) == 0 && o == rE, "ext after fault");
        expect(loc(eng, L, o) == 0 && o == rL, "local after fault");
    }
    std::puts("step: serving loop, 4 slots, two submitters");
    std::fflush(stdout);
    {
        const int64_t gl = 2000000;
        uint8_t *ref = wl_reference(91, gl, 200, 0.10);
        std::vector<uint8_t> pac((size_t)(gl + 3) / 4);
        wl_pack_pac(ref, gl, pac.data());
        const hipsw_scoring sc = scoring();
        hipsw_server *srv = hipsw_server_create(eng, 4, (int64_t)4 << 30, &sc, 1, 4, 512);
        expect(srv && hipsw_server_set_reference(srv, pac.data(), gl) == 0, "server create");
        std::vector<ChunkView> cs;
        for (int k = 0; k < 8; ++k) cs.push_back(chunk(ref, gl, 500 + k, 500 + 300 * k));
        std::vector<ChunkOut> alone(cs.size()), mix(cs.size());
        for (size_t k = 0; k < cs.size(); ++k)
            wait_chunk(srv, hipsw_server_submit(srv, &cs[k].c), cs[k].c, alone[k]);
        std::vector<int64_t> tickets(cs.size(), 0);
        std::vector<std::thread> sub;
        for (int s = 0; s < 2; ++s)
            sub.emplace_back([&, s] {
                for (size_t k = s; k < cs.size(); k += 2) tickets[k] = hipsw_server_submit(srv, &cs[k].c);
            });
        for (auto &x : sub) x.join();
        std::thread waiter([&] {
            for (size_t k = 0; k < cs.size(); ++k) wait_chunk(srv, tickets[k], cs[k].c, mix[k]);
        });
        waiter.join();
        for (size_t k = 0; k < cs.size(); ++k)
            expect(alone[k].status == 0 && mix[k].status == 0 && same(alone[k], mix[k]), "chunk alone == mix");
        hipsw_fault_inject(eng, HIPSW_FAULT_LAUNCH, 1);
        ChunkOut f;
        wait_chunk(srv, hipsw_server_submit(srv, &cs[0].c), cs[0].c, f);
        expect(f.status == HIPSW_E_DEVICE, "chunk fault status");
        hipsw_fault_inject(eng, HIPSW_FAULT_NONE, 0);
        wait_chunk(srv, hipsw_server_submit(srv, &cs[1].c), cs[1].c, f);
        expect(f.status == 0 && same(f, alone[1]), "chunk after fault");
        hipsw_server_destroy(srv);
        for (auto &c : cs) wl_chunk_free(c.h);
        wl_free_bytes(ref);
    }
    for (Jobs *j : {&E, &U, &N, &L}) wl_free(j->h);
    hipsw_engine_destroy(eng);
    std::printf("%s: %d mismatches\n", g_bad ? "SANITIZE DRIVER FAILED" : "SANITIZE DRIVER OK",
                g_bad.load());
    std::fflush(stdout);   // the verdict is out before the runtime's own teardown
    return g_bad ? 1 : 0;
}
