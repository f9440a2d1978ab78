// On-disk geometry cache: sal_geom_save writes everything a geometry handle derived from its points
// (the tree, the DFS layout, the DTT list, the STT lists), sal_geom_load rebuilds an equivalent
// handle from the file without re-running the tree build or the traversals.  The reference's
// drivers rebuild the tree and lists on every run (code.f90:72-82, mom6_code.f90:177-220); a
// serving process that restarts reads them back instead.
//
// File layout (little-endian, the host's own byte order is checked against the header):
//   magic "SALGEOM\0", u32 format version, u32 byte-order mark 0x01020304
//   sections, each: u32 tag, u32 element size in bytes, u64 element count, the elements
//   a final section with tag END and the FNV-1a 64 hash of every byte before it
// The loader checks the magic, version, byte order, every section's tag, element size and count
// against the header values, and the hash; any mismatch fails with SAL_ERR_IO and no handle.
// Evaluation plans and stored operators are not written: they are rebuilt on first use from the
// lists, with the same results.

#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstring>
#include <memory>

#include "common.h"

namespace sal {

namespace {

constexpr char MAGIC[8] = {'S', 'A', 'L', 'G', 'E', 'O', 'M', '\0'};
constexpr uint32_t VERSION = 1, BOM = 0x01020304u;

enum Tag : uint32_t {
    T_HEADER = 1, T_X, T_Y, T_Z, T_XI, T_ETA, T_PFACE, T_ANC,
    T_LEVEL, T_LEAF, T_PARENT, T_CHILD1, T_FACE, T_COUNT, T_LEVEL_LO, T_BOUNDS, T_RADIUS, T_CENTRE, T_START,
    T_PERM, T_IPERM, T_PX, T_PY, T_PZ,
    T_IT, T_IS, T_TY,
    T_PC_T, T_PC_S, T_PP_T, T_PP_S, T_STT_COUNTS, T_STT_TX, T_STT_TY, T_STT_TZ, T_STT_PC_OFF, T_STT_PP_OFF,
    T_END = 0xE0Du
};

// Scalars of the handle, one fixed-size record
struct Header {
    int64_t n, nint, npc, npp, stt_nt, point_ties, pair_ties;
    int64_t type_count[4];
    int32_t thresh, nlev, P, list_thresh, n_level_lo, pad;
};

struct Fnv {
    uint64_t h = 1469598103934665603ull;
    void add(const void* p, size_t b) {
        const unsigned char* c = (const unsigned char*)p;
        for (size_t i = 0; i < b; ++i) h = (h ^ c[i]) * 1099511628211ull;
    }
};

struct File {
    FILE* f = nullptr;
    ~File() { if (f) fclose(f); }
};

class Writer {
   public:
    explicit Writer(FILE* f) : f_(f) {}
    void raw(const void* p, size_t b) {
        if (b && fwrite(p, 1, b, f_) != b) throw Error{SAL_ERR_IO, "geometry cache: write failed"};
        h_.add(p, b);
    }
    template <class T>
    void host(uint32_t tag, const T* p, size_t count) {
        const uint32_t es = sizeof(T);
        const uint64_t c = count;
        raw(&tag, 4);
        raw(&es, 4);
        raw(&c, 8);
        raw(p, count * sizeof(T));
    }
    template <class T>
    void dev(uint32_t tag, const T* d, size_t count, hipStream_t s) {
        std::vector<T> h(count);
        d2h(h.data(), d, count, s);
        SAL_HIP(hipStreamSynchronize(s));
        host(tag, h.data(), count);
    }
    void end() {
        const uint32_t tag = T_END, es = 8;
        const uint64_t c = 1;
        raw(&tag, 4);
        raw(&es, 4);
        raw(&c, 8);
        const uint64_t hash = h_.h;
        if (fwrite(&hash, 1, 8, f_) != 8) throw Error{SAL_ERR_IO, "geometry cache: write failed"};
    }

   private:
    FILE* f_;
    Fnv h_;
};

class Reader {
   public:
    explicit Reader(FILE* f) : f_(f) {}
    void raw(void* p, size_t b) {
        if (b && fread(p, 1, b, f_) != b) throw Error{SAL_ERR_IO, "geometry cache: file is truncated"};
        h_.add(p, b);
    }
    // reads the next section, which must be `tag` with `count` elements of T
    template <class T>
    void host(uint32_t tag, T* p, size_t count) {
        uint32_t t = 0, es = 0;
        uint64_t c = 0;
        raw(&t, 4);
        raw(&es, 4);
        raw(&c, 8);
        if (t != tag || es != sizeof(T) || c != count)
            throw Error{SAL_ERR_IO, "geometry cache: section " + std::to_string(tag) + " is missing or has the wrong size"};
        raw(p, count * sizeof(T));
    }
    template <class T>
    void dev(uint32_t tag, DBuf<T>& d, size_t count, size_t alloc_min, hipStream_t s) {
        std::vector<T> h(count);
        host(tag, h.data(), count);
        d.alloc(std::max(count, alloc_min));
        h2d(d.p, h.data(), count, s);
        SAL_HIP(hipStreamSynchronize(s));
    }
    void end() {
        uint32_t t = 0, es = 0;
        uint64_t c = 0, hash = 0;
        raw(&t, 4);
        raw(&es, 4);
        raw(&c, 8);
        const uint64_t want = h_.h;
        if (t != T_END || es != 8 || c != 1 || fread(&hash, 1, 8, f_) != 8 || hash != want)
            throw Error{SAL_ERR_IO, "geometry cache: checksum mismatch (file damaged or not a geometry cache)"};
        char extra;
        if (fread(&extra, 1, 1, f_) != 0) throw Error{SAL_ERR_IO, "geometry cache: trailing bytes after the end"};
    }

   private:
    FILE* f_;
    Fnv h_;
};

}  // namespace

static void save(sal_geom g, const char* path) {
    hipStream_t s = g->stream;
    SAL_HIP(hipStreamSynchronize(s));
    SttState st;
    stt_get(g, st);
    Header h{};
    h.n = g->n;
    h.nint = g->nint;
    h.npc = g->npc;
    h.npp = g->npp;
    h.stt_nt = st.nt;
    h.point_ties = g->point_ties;
    h.pair_ties = g->pair_ties;
    for (int k = 0; k < 4; ++k) h.type_count[k] = g->type_count[k];
    h.thresh = g->thresh;
    h.nlev = g->nlev;
    h.P = g->P;
    h.list_thresh = g->list_thresh;
    h.n_level_lo = (int32_t)g->level_lo.size();
    const std::string tmp = std::string(path) + ".tmp";
    {
        File fh;
        fh.f = fopen(tmp.c_str(), "wb");
        SAL_REQUIRE(fh.f, SAL_ERR_IO, std::string("geometry cache: cannot create ") + tmp);
        Writer w(fh.f);
        const uint32_t ver = VERSION, bom = BOM;
        w.raw(MAGIC, 8);
        w.raw(&ver, 4);
        w.raw(&bom, 4);
        w.host(T_HEADER, &h, 1);
        const size_t n = (size_t)g->n, P = (size_t)g->P;
        w.dev(T_X, g->x.p, n, s);
        w.dev(T_Y, g->y.p, n, s);
        w.dev(T_Z, g->z.p, n, s);
        w.dev(T_XI, g->xi.p, n, s);
        w.dev(T_ETA, g->eta.p, n, s);
        w.dev(T_PFACE, g->pface.p, n, s);
        w.dev(T_ANC, g->anc.p, (size_t)g->nlev * n, s);
        w.host(T_LEVEL, g->level.data(), P);
        w.host(T_LEAF, g->is_leaf.data(), P);
        w.host(T_PARENT, g->parent.data(), P);
        w.host(T_CHILD1, g->child1.data(), P);
        w.host(T_FACE, g->face.data(), P);
        w.host(T_COUNT, g->count.data(), P);
        w.host(T_LEVEL_LO, g->level_lo.data(), g->level_lo.size());
        w.host(T_BOUNDS, g->bounds.data(), 6 * P);
        w.host(T_RADIUS, g->radius.data(), P);
        w.host(T_CENTRE, g->centre.data(), 3 * P);
        w.host(T_START, g->start.data(), P);
        w.dev(T_PERM, g->perm.p, n, s);
        w.dev(T_IPERM, g->iperm.p, n, s);
        w.dev(T_PX, g->px.p, n, s);
        w.dev(T_PY, g->py.p, n, s);
        w.dev(T_PZ, g->pz.p, n, s);
        w.dev(T_IT, g->it.p, (size_t)g->nint, s);
        w.dev(T_IS, g->is.p, (size_t)g->nint, s);
        w.dev(T_TY, g->ty.p, (size_t)g->nint, s);
        w.dev(T_PC_T, g->pc_t.p, (size_t)g->npc, s);
        w.dev(T_PC_S, g->pc_s.p, (size_t)g->npc, s);
        w.dev(T_PP_T, g->pp_t.p, (size_t)g->npp, s);
        w.dev(T_PP_S, g->pp_s.p, (size_t)g->npp, s);
        w.dev(T_STT_COUNTS, g->stt_counts.p, 2 * (size_t)st.nt, s);
        w.host(T_STT_TX, st.tx.data(), (size_t)st.nt);
        w.host(T_STT_TY, st.ty.data(), (size_t)st.nt);
        w.host(T_STT_TZ, st.tz.data(), (size_t)st.nt);
        w.host(T_STT_PC_OFF, st.pc_off.data(), (size_t)st.nt);
        w.host(T_STT_PP_OFF, st.pp_off.data(), (size_t)st.nt);
        w.end();
        SAL_REQUIRE(fflush(fh.f) == 0, SAL_ERR_IO, "geometry cache: write failed");
    }
    SAL_REQUIRE(rename(tmp.c_str(), path) == 0, SAL_ERR_IO, std::string("geometry cache: cannot rename to ") + path);
}

static sal_geom load(const char* path) {
    File fh;
    fh.f = fopen(path, "rb");
    SAL_REQUIRE(fh.f, SAL_ERR_IO, std::string("geometry cache: cannot open ") + path);
    Reader r(fh.f);
    char magic[8];
    uint32_t ver = 0, bom = 0;
    r.raw(magic, 8);
    SAL_REQUIRE(memcmp(magic, MAGIC, 8) == 0, SAL_ERR_IO, "geometry cache: not a geometry cache file");
    r.raw(&ver, 4);
    SAL_REQUIRE(ver == VERSION, SAL_ERR_IO, "geometry cache: format version " + std::to_string(ver) +
                                                " (this library reads " + std::to_string(VERSION) + ")");
    r.raw(&bom, 4);
    SAL_REQUIRE(bom == BOM, SAL_ERR_IO, "geometry cache: written on a host of the other byte order");
    Header h{};
    r.host(T_HEADER, &h, 1);
    SAL_REQUIRE(h.n >= 1 && h.n < INT32_MAX && h.P >= 6 && h.nlev >= 1 && h.n_level_lo == h.nlev + 1 &&
                    h.thresh >= 1 && h.nint >= 0 && h.nint < INT32_MAX && h.npc >= 0 && h.npp >= 0 &&
                    h.stt_nt >= 0 && h.stt_nt < INT32_MAX,
                SAL_ERR_IO, "geometry cache: header values out of range");
    auto g = std::make_unique<sal_geom_s>();
    SAL_HIP(hipStreamCreateWithFlags(&g->stream, hipStreamNonBlocking));
    hipStream_t s = g->stream;
    const size_t n = (size_t)h.n, P = (size_t)h.P;
    g->n = h.n;
    g->thresh = h.thresh;
    g->nlev = h.nlev;
    g->P = h.P;
    g->list_thresh = h.list_thresh;
    g->nint = h.nint;
    g->npc = h.npc;
    g->npp = h.npp;
    g->point_ties = h.point_ties;
    g->pair_ties = h.pair_ties;
    for (int k = 0; k < 4; ++k) g->type_count[k] = h.type_count[k];
    auto hv = [&](uint32_t tag, auto& v, size_t count) {
        v.resize(count);
        r.host(tag, v.data(), count);
    };
    // index arrays are read to the host and range-checked before they reach the device, so that no
    // kernel indexes out of bounds on a file whose contents are inconsistent
    auto up = [&](auto& d, const auto& v) {
        d.alloc(std::max<size_t>(v.size(), 1));
        h2d(d.p, v.data(), v.size(), s);
        SAL_HIP(hipStreamSynchronize(s));
    };
    auto in_range = [](const std::vector<int>& v, int lo, int64_t hi) {
        for (int a : v)
            if (a < lo || a >= hi) return false;
        return true;
    };
    r.dev(T_X, g->x, n, 1, s);
    r.dev(T_Y, g->y, n, 1, s);
    r.dev(T_Z, g->z, n, 1, s);
    r.dev(T_XI, g->xi, n, 1, s);
    r.dev(T_ETA, g->eta, n, 1, s);
    {
        std::vector<int> v;
        hv(T_PFACE, v, n);
        SAL_REQUIRE(in_range(v, 1, 7), SAL_ERR_IO, "geometry cache: point face out of range");
        up(g->pface, v);
        hv(T_ANC, v, (size_t)h.nlev * n);
        SAL_REQUIRE(in_range(v, -1, h.P), SAL_ERR_IO, "geometry cache: point panel out of range");
        up(g->anc, v);
    }
    hv(T_LEVEL, g->level, P);
    hv(T_LEAF, g->is_leaf, P);
    hv(T_PARENT, g->parent, P);
    hv(T_CHILD1, g->child1, P);
    hv(T_FACE, g->face, P);
    hv(T_COUNT, g->count, P);
    hv(T_LEVEL_LO, g->level_lo, (size_t)h.n_level_lo);
    hv(T_BOUNDS, g->bounds, 6 * P);
    hv(T_RADIUS, g->radius, P);
    hv(T_CENTRE, g->centre, 3 * P);
    hv(T_START, g->start, P);
    // panel arrays index one another and the point ranges: check them before any kernel uses them
    SAL_REQUIRE(g->level_lo[0] == 0 && g->level_lo.back() == h.P, SAL_ERR_IO, "geometry cache: bad level ranges");
    for (size_t i = 0; i < P; ++i) {
        const bool ok = g->parent[i] >= -1 && g->parent[i] < h.P && g->child1[i] >= -1 && g->child1[i] + 4 <= h.P &&
                        g->count[i] >= 0 && g->start[i] >= 0 && g->start[i] + g->count[i] <= h.n &&
                        g->face[i] >= 1 && g->face[i] <= 6 && g->level[i] >= 0 && g->level[i] < h.nlev;
        SAL_REQUIRE(ok, SAL_ERR_IO, "geometry cache: panel " + std::to_string(i) + " is inconsistent");
    }
    // The tree's structure, which the plans and kernels rely on (a checksum only detects damage, not a
    // consistently written wrong file): six roots at level 0 without parents, splitting the points in
    // order; each level's panels one id range; a leaf has no children and fewer than thresh points; a
    // non-leaf's four children are its own (parent, next level) and split its point range in child
    // order; every other panel is a child of its parent.
    SAL_REQUIRE(h.P >= 6, SAL_ERR_IO, "geometry cache: fewer than six root panels");
    {
        int64_t acc = 0;
        for (int i = 0; i < 6; ++i) {
            SAL_REQUIRE(g->level[i] == 0 && g->parent[i] == -1 && g->start[i] == acc, SAL_ERR_IO,
                        "geometry cache: root panel " + std::to_string(i) + " is inconsistent");
            acc += g->count[i];
        }
        SAL_REQUIRE(acc == h.n, SAL_ERR_IO, "geometry cache: the root panels do not hold every point");
    }
    for (int L = 0; L + 1 < h.n_level_lo; ++L) {
        SAL_REQUIRE(g->level_lo[L] <= g->level_lo[L + 1], SAL_ERR_IO, "geometry cache: level ranges decrease");
        for (int A = g->level_lo[L]; A < g->level_lo[L + 1]; ++A)
            SAL_REQUIRE(g->level[A] == L, SAL_ERR_IO,
                        "geometry cache: panel " + std::to_string(A) + " lies outside its level's range");
    }
    for (size_t i = 0; i < P; ++i) {
        const int c = g->child1[i];
        bool ok = (g->is_leaf[i] == 0 || g->is_leaf[i] == 1) && (g->is_leaf[i] == 1) == (c < 0);
        if (ok && c < 0) {
            ok = g->count[i] < h.thresh;
        } else if (ok) {
            int64_t a = g->start[i], sum = 0;
            for (int k = 0; ok && k < 4; ++k) {
                ok = g->parent[c + k] == (int)i && g->level[c + k] == g->level[i] + 1 && g->start[c + k] == a;
                a += g->count[c + k];
                sum += g->count[c + k];
            }
            ok = ok && sum == g->count[i];
        }
        const int p = g->parent[i];
        if (ok && i >= 6) ok = p >= 0 && g->child1[p] >= 0 && (int)i >= g->child1[p] && (int)i < g->child1[p] + 4;
        SAL_REQUIRE(ok, SAL_ERR_IO, "geometry cache: panel " + std::to_string(i) + " breaks the tree's structure");
    }
    {
        std::vector<int> pm, ip;
        hv(T_PERM, pm, n);
        hv(T_IPERM, ip, n);
        bool ok = in_range(pm, 0, h.n) && in_range(ip, 0, h.n);
        for (size_t k = 0; ok && k < n; ++k) ok = ip[pm[k]] == (int)k;
        SAL_REQUIRE(ok, SAL_ERR_IO, "geometry cache: the DFS order is not a permutation");
        up(g->perm, pm);
        up(g->iperm, ip);
    }
    upload_panels(g.get());
    r.dev(T_PX, g->px, n, 1, s);
    r.dev(T_PY, g->py, n, 1, s);
    r.dev(T_PZ, g->pz, n, 1, s);
    const size_t ni = (size_t)h.nint;
    {
        std::vector<int> it(ni), is(ni), ty(ni);
        r.host(T_IT, it.data(), ni);
        r.host(T_IS, is.data(), ni);
        r.host(T_TY, ty.data(), ni);
        int64_t tc[4] = {0, 0, 0, 0};
        for (size_t k = 0; k < ni; ++k) {
            SAL_REQUIRE(it[k] >= 0 && it[k] < h.P && is[k] >= 0 && is[k] < h.P && ty[k] >= 0 && ty[k] <= 3,
                        SAL_ERR_IO, "geometry cache: interaction list entry out of range");
            tc[ty[k]]++;
        }
        for (int k = 0; k < 4; ++k)
            SAL_REQUIRE(tc[k] == h.type_count[k], SAL_ERR_IO, "geometry cache: interaction type counts differ");
        g->it.alloc(std::max<size_t>(ni, 1));
        g->is.alloc(std::max<size_t>(ni, 1));
        g->ty.alloc(std::max<size_t>(ni, 1));
        h2d(g->it.p, it.data(), ni, s);
        h2d(g->is.p, is.data(), ni, s);
        h2d(g->ty.p, ty.data(), ni, s);
        SAL_HIP(hipStreamSynchronize(s));
    }
    const size_t nt = (size_t)h.stt_nt;
    std::vector<int> pct, pcs, ppt, pps, cnts;
    hv(T_PC_T, pct, (size_t)h.npc);
    hv(T_PC_S, pcs, (size_t)h.npc);
    hv(T_PP_T, ppt, (size_t)h.npp);
    hv(T_PP_S, pps, (size_t)h.npp);
    hv(T_STT_COUNTS, cnts, 2 * nt);
    SttState st;
    st.nt = h.stt_nt;
    hv(T_STT_TX, st.tx, nt);
    hv(T_STT_TY, st.ty, nt);
    hv(T_STT_TZ, st.tz, nt);
    hv(T_STT_PC_OFF, st.pc_off, nt);
    hv(T_STT_PP_OFF, st.pp_off, nt);
    r.end();
    bool ok = in_range(pct, 0, h.stt_nt) && in_range(ppt, 0, h.stt_nt) && in_range(pcs, 0, h.P) &&
              in_range(pps, 0, h.P) && (nt > 0 || (h.npc == 0 && h.npp == 0));
    for (size_t i = 0; ok && i < nt; ++i)
        ok = cnts[i] >= 0 && cnts[nt + i] >= 0 && st.pc_off[i] >= 0 && st.pp_off[i] >= 0 &&
             (int64_t)st.pc_off[i] + cnts[i] <= h.npc && (int64_t)st.pp_off[i] + cnts[nt + i] <= h.npp;
    SAL_REQUIRE(ok, SAL_ERR_IO, "geometry cache: single tree traversal lists out of range");
    up(g->pc_t, pct);
    up(g->pc_s, pcs);
    up(g->pp_t, ppt);
    up(g->pp_s, pps);
    if (nt) up(g->stt_counts, cnts);
    stt_set(g.get(), st);
    // every device array the kernels read is in place (the lists may be empty but are allocated)
    const void* need[] = {g->x.p, g->y.p, g->z.p, g->xi.p, g->eta.p, g->pface.p, g->anc.p, g->d_level.p,
                          g->d_is_leaf.p, g->d_parent.p, g->d_child1.p, g->d_face.p, g->d_count.p, g->d_bounds.p,
                          g->d_radius.p, g->d_centre.p, g->d_start.p, g->perm.p, g->iperm.p, g->px.p, g->py.p,
                          g->pz.p, g->it.p, g->is.p, g->ty.p, g->pc_t.p, g->pc_s.p, g->pp_t.p, g->pp_s.p};
    for (const void* p : need) SAL_REQUIRE(p != nullptr, SAL_ERR_STATE, "geometry cache: a device array was not restored");
    return g.release();
}

}  // namespace sal

using namespace sal;

extern "C" {

int sal_geom_save(sal_geom g, const char* path) {
    return guard([&] {
        SAL_REQUIRE(g && path && *path, SAL_ERR_ARG, "null handle or path");
        save(g, path);
    });
}

int sal_geom_load(const char* path, sal_geom* out) {
    return guard([&] {
        SAL_REQUIRE(path && *path && out, SAL_ERR_ARG, "null pointer");
        *out = nullptr;
        *out = load(path);
    });
}

}  // extern "C"
