// Result and checkpoint files (layout in qk_store.hpp).
#include "qk_store.hpp"

#include "qk_threads.hpp"

#include <cstdio>
#include <cstring>
#include <memory>

namespace qk {

namespace {

constexpr char kMagic[8] = {'Q', 'K', 'R', 'Y', 'L', 'O', 'V', '\0'};
constexpr u32 kVersion = 1;
constexpr u32 kKindEigen = 1, kKindLanczos = 2;
constexpr u64 kP = 0x100000001b3ull;

struct Header {
    char magic[8];
    u32 version, kind;
    u64 dim, fingerprint, a, b, c;
    double e;
};
static_assert(sizeof(Header) == 64, "file header layout");

// Checksum over 8-byte words: the stream is cut into blocks of 2^17 words
// (1 MiB), each block hashed on its own (h = n; h = (h ^ w) * P, h ^= h >> 29
// per word), and the block hashes folded in order (t = (t ^ h) * P,
// t ^= t >> 31).  Blocks are independent, so long spans hash on all host
// threads.
class BlockHash {
public:
    static constexpr size_t kBlock = size_t(1) << 17;
    void add(const u64* w, size_t n) {
        while (n && !carry_.empty()) {
            const size_t take = std::min(n, kBlock - carry_.size());
            carry_.insert(carry_.end(), w, w + take);
            w += take;
            n -= take;
            if (carry_.size() == kBlock) {
                fold(hash_words(carry_.data(), kBlock));
                carry_.clear();
            }
        }
        const size_t nfull = n / kBlock;
        if (nfull) {
            std::vector<u64> hs(nfull);
            run_ranges(nfull, host_threads(nfull), [&](u64 a, u64 b) {
                for (u64 i = a; i < b; ++i) hs[i] = hash_words(w + i * kBlock, kBlock);
            });
            for (u64 h : hs) fold(h);
            w += nfull * kBlock;
            n -= nfull * kBlock;
        }
        if (n) carry_.assign(w, w + n);
    }
    u64 finish() {
        if (!carry_.empty()) fold(hash_words(carry_.data(), carry_.size()));
        carry_.clear();
        return t_;
    }

private:
    static u64 hash_words(const u64* w, size_t n) {
        u64 h = 0xcbf29ce484222325ull ^ u64(n);
        for (size_t i = 0; i < n; ++i) {
            h = (h ^ w[i]) * kP;
            h ^= h >> 29;
        }
        return h;
    }
    void fold(u64 h) {
        t_ = (t_ ^ h) * kP;
        t_ ^= t_ >> 31;
    }
    u64 t_ = 0x84222325cbf29ce4ull;
    std::vector<u64> carry_;
};

struct File {
    std::FILE* f = nullptr;
    std::string path;
    File(const std::string& p, const char* mode) : path(p) {
        f = std::fopen(p.c_str(), mode);
        if (!f) fail("cannot open " + p);
    }
    ~File() {
        if (f) std::fclose(f);
    }
    void write(const void* p, size_t bytes) {
        if (bytes && std::fwrite(p, 1, bytes, f) != bytes) fail("write failed: " + path);
    }
    void read(void* p, size_t bytes) {
        if (bytes && std::fread(p, 1, bytes, f) != bytes) fail("truncated file: " + path);
    }
    void close() {
        if (f && std::fclose(f) != 0) {
            f = nullptr;
            fail("write failed: " + path);
        }
        f = nullptr;
    }
};

// hashes and writes 8-byte words
struct Writer {
    File file;
    BlockHash hash;
    explicit Writer(const std::string& p) : file(p, "wb") {}
    void put(const void* p, size_t bytes) {
        hash.add(static_cast<const u64*>(p), bytes / 8);
        file.write(p, bytes);
    }
    void finish() {
        const u64 h = hash.finish();
        file.write(&h, sizeof h);
        file.close();
    }
};

struct Reader {
    File file;
    BlockHash hash;
    explicit Reader(const std::string& p) : file(p, "rb") {}
    void get(void* p, size_t bytes) {
        file.read(p, bytes);
        hash.add(static_cast<const u64*>(p), bytes / 8);
    }
    void check() {
        u64 h = 0;
        file.read(&h, sizeof h);
        if (h != hash.finish()) fail("checksum mismatch: " + file.path);
    }
};

Header read_header(Reader& r, u32 kind) {
    Header h{};
    r.get(&h, sizeof h);
    if (std::memcmp(h.magic, kMagic, 8) != 0) fail("not a qkrylov file: " + r.file.path);
    if (h.version != kVersion) fail("unsupported qkrylov file version: " + r.file.path);
    if (h.kind != kind) fail("wrong qkrylov file kind: " + r.file.path);
    return h;
}

Header make_header(u32 kind, u64 dim, u64 fp) {
    Header h{};
    std::memcpy(h.magic, kMagic, 8);
    h.version = kVersion;
    h.kind = kind;
    h.dim = dim;
    h.fingerprint = fp;
    return h;
}

// a pinned staging buffer for device-to-host copies of long vectors
struct Staging {
    void* p = nullptr;
    size_t bytes;
    explicit Staging(size_t b) : bytes(b) { QK_HIP(hipHostMalloc(&p, b)); }
    ~Staging() { (void)hipHostFree(p); }
};

}  // namespace

u64 hamiltonian_fingerprint(const Hamiltonian& H) {
    BlockHash h;
    std::vector<u64> w;
    const Basis& B = *H.basis;
    const Sector& s = B.sector();
    w = {u64(H.site), u64(B.nsites()), u64(B.kind()), B.size(),
         u64(s.use_sz), u64(u32(s.sz2)), u64(s.use_nup), u64(s.use_ndn), u64(u32(s.nup)), u64(u32(s.ndn)),
         u64(s.use_n), u64(u32(s.n)), u64(s.use_nb), u64(u32(s.nb)), u64(H.ops.terms.size())};
    for (const OpTerm& t : H.ops.terms) {
        u64 re, im;
        std::memcpy(&re, &t.coeff.re, 8);
        std::memcpy(&im, &t.coeff.im, 8);
        w.push_back(re);
        w.push_back(im);
        w.push_back(u64(t.factors.size()));
        for (const OpFactor& f : t.factors) {
            u64 name = 0;
            std::memcpy(&name, f.op.data(), std::min<size_t>(8, f.op.size()));
            w.push_back(name);
            w.push_back(u64(f.op.size()));
            w.push_back(u64(u32(f.site)));
        }
    }
    h.add(w.data(), w.size());
    return h.finish();
}

void save_lanczos_checkpoint(const std::string& path, const Hamiltonian& H, const LanczosState& st,
                             const cplx* const* dev, hipStream_t stream) {
    const u64 d = H.dim();
    if (st.next_step < 0 || st.alphas.size() != size_t(st.next_step) ||
        st.betas.size() != size_t(st.next_step))
        fail("lanczos checkpoint: inconsistent state");
    Header h = make_header(kKindLanczos, d, hamiltonian_fingerprint(H));
    h.a = u64(st.next_step);
    h.b = st.alphas.size();
    h.c = st.betas.size();
    h.e = st.e_old;
    // written under a temporary name and renamed, so an interrupted save
    // leaves the previous checkpoint intact
    const std::string tmp = path + ".part";
    {
        Writer out(tmp);
        out.put(&h, sizeof h);
        out.put(st.alphas.data(), sizeof(double) * st.alphas.size());
        out.put(st.betas.data(), sizeof(double) * st.betas.size());
        const size_t chunk = std::min<size_t>(size_t(64) << 20, sizeof(cplx) * d);
        Staging stage(chunk);
        for (int j = 0; j <= st.next_step; ++j) {
            const char* src = reinterpret_cast<const char*>(dev[j]);
            for (size_t off = 0; off < sizeof(cplx) * d; off += chunk) {
                const size_t nb = std::min(chunk, sizeof(cplx) * d - off);
                QK_HIP(hipMemcpyAsync(stage.p, src + off, nb, hipMemcpyDeviceToHost, stream));
                QK_HIP(hipStreamSynchronize(stream));
                out.put(stage.p, nb);
            }
        }
        out.finish();
    }
    if (std::rename(tmp.c_str(), path.c_str()) != 0) fail("cannot rename " + tmp + " to " + path);
}

LanczosState load_lanczos_checkpoint(const std::string& path, const Hamiltonian& H) {
    Reader in(path);
    const Header h = read_header(in, kKindLanczos);
    if (h.dim != H.dim()) fail("lanczos checkpoint: dimension differs from the Hamiltonian's");
    if (h.fingerprint != hamiltonian_fingerprint(H))
        fail("lanczos checkpoint: written for a different Hamiltonian");
    if (h.b != h.a || h.c != h.a || h.a > u64(1) << 30) fail("lanczos checkpoint: inconsistent header");
    LanczosState st;
    st.next_step = int(h.a);
    st.e_old = h.e;
    st.alphas.resize(h.b);
    st.betas.resize(h.c);
    in.get(st.alphas.data(), sizeof(double) * h.b);
    in.get(st.betas.data(), sizeof(double) * h.c);
    st.vectors.resize(size_t(h.a + 1) * h.dim);
    in.get(st.vectors.data(), sizeof(cplx) * st.vectors.size());
    in.check();
    return st;
}

void save_eigenpairs(const std::string& path, u64 dim, u64 fingerprint, int n, const double* values,
                     const cplx* vectors) {
    if (n < 0 || (n > 0 && !values)) fail("eigenpairs: invalid arguments");
    Header h = make_header(kKindEigen, dim, fingerprint);
    h.a = u64(n);
    h.b = vectors ? 1 : 0;
    const std::string tmp = path + ".part";
    {
        Writer out(tmp);
        out.put(&h, sizeof h);
        out.put(values, sizeof(double) * size_t(n));
        if (vectors) out.put(vectors, sizeof(cplx) * size_t(n) * dim);
        out.finish();
    }
    if (std::rename(tmp.c_str(), path.c_str()) != 0) fail("cannot rename " + tmp + " to " + path);
}

EigenFile load_eigenpairs(const std::string& path, bool want_vectors) {
    Reader in(path);
    const Header h = read_header(in, kKindEigen);
    if (h.a > u64(1) << 30 || h.b > 1) fail("eigenpairs: inconsistent header");
    EigenFile e;
    e.dim = h.dim;
    e.fingerprint = h.fingerprint;
    e.has_vectors = h.b == 1;
    e.values.resize(h.a);
    in.get(e.values.data(), sizeof(double) * h.a);
    if (!want_vectors) return e;   // header and values only: no checksum
    if (e.has_vectors) {
        e.vectors.resize(size_t(h.a) * h.dim);
        in.get(e.vectors.data(), sizeof(cplx) * e.vectors.size());
    }
    in.check();
    return e;
}

}  // namespace qk
