/*
 * Checkpoint files of GPU state in the reference's layout
 * (nbody/src/nbody_checkpoint.c:76-104 header, nbFreezeState :488-536,
 * nbThawState :396-486, nbVerifyCheckpointHeader :132-182, file size
 * :213-214), for x86-64 hosts with real = double:
 *
 *   header   char[128] "mwnbody"; uint32 majorVersion, minorVersion, nbody,
 *            step, realSize, ptrSize, nOrbitTrace, nShiftLMC, treeIncest;
 *            double rsize (offset 168); NBodyCtx ctx (offset 176, the
 *            struct is 16-byte aligned, so the header is 176 + ctx rounded
 *            up to 16 bytes)
 *   size_t   sizeof(NBodyState), then the raw NBodyState
 *   Body[n]  bodytab; Body[n] bestLikelihoodBodyTab (96 B each: pos.xyzw
 *            at 0, next at 32, mass at 40, short type at 48, uint id at 52,
 *            vel.xyzw at 64)
 *   mwvector[nOrbitTrace] orbit trace
 *   if nShiftLMC: mwvector[nShiftLMC] shifts, LMCpos, LMCvel
 *   "end\0"
 *
 * and the file is hdr + 4 + 2 n 96 + 32 (nOrbitTrace + nShiftLMC) + 64 +
 * sizeof(NBodyState) + 8 bytes long whether or not the LMC part is written
 * (the writer sizes the file before filling it; the rest stays zero).
 *
 * The NBodyCtx and NBodyState images belong to the host program: they are
 * written and returned byte for byte.  The pack / unpack calls are host-only
 * (no device); nbhip_checkpoint_write / _read move the bodies between the
 * file and a GPU state and carry its step, root size and incest flag.
 */
#include "nbhip_internal.h"

#include <cerrno>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <unistd.h>
#include <vector>

namespace {

const char kMagic[] = "mwnbody";
const char kTail[] = "end";
const size_t kBody = 96;          /* sizeof(Body) */
const size_t kVec = 32;           /* sizeof(mwvector) */
const size_t kHdrFixed = 176;     /* offset of ctx inside the header */

size_t round16(size_t v) { return (v + 15) & ~(size_t) 15; }

size_t header_size(size_t ctx_size) { return round16(kHdrFixed + ctx_size); }

size_t file_size(size_t ctx_size, size_t state_size, uint64_t nbody, uint64_t ntrace,
                 uint64_t nshift)
{
    return header_size(ctx_size) + sizeof(kTail) + 2 * nbody * kBody + ntrace * kVec +
           nshift * kVec + 2 * kVec + state_size + sizeof(uint64_t);
}

void put_u32(unsigned char* p, uint32_t v) { memcpy(p, &v, 4); }
uint32_t get_u32(const unsigned char* p)
{
    uint32_t v;
    memcpy(&v, p, 4);
    return v;
}
void put_f64(unsigned char* p, double v) { memcpy(p, &v, 8); }
double get_f64(const unsigned char* p)
{
    double v;
    memcpy(&v, p, 8);
    return v;
}

/* One Body record: pos (w = 0), next = NULL, mass, type, id, vel (w = 0). */
void put_body(unsigned char* p, const double* x, const double* v, double m, short type,
              unsigned id)
{
    memset(p, 0, kBody);
    for (int d = 0; d < 3; ++d) put_f64(p + 8 * d, x[d]);
    put_f64(p + 40, m);
    memcpy(p + 48, &type, sizeof(type));
    put_u32(p + 52, id);
    for (int d = 0; d < 3; ++d) put_f64(p + 64 + 8 * d, v[d]);
}

void get_body(const unsigned char* p, double* x, double* v, double* m, short* type,
              unsigned* id)
{
    for (int d = 0; d < 3; ++d) {
        if (x) x[d] = get_f64(p + 8 * d);
        if (v) v[d] = get_f64(p + 64 + 8 * d);
    }
    if (m) *m = get_f64(p + 40);
    if (type) memcpy(type, p + 48, sizeof(short));
    if (id) *id = get_u32(p + 52);
}

bool write_all(FILE* f, const void* p, size_t n) { return n == 0 || fwrite(p, 1, n, f) == n; }

}  // namespace

extern "C" unsigned nbhip_checkpoint_pack(const char* path, const nbhip_checkpoint_header* hdr,
                                          const nbhip_checkpoint_io* io, const double* pos,
                                          const double* vel, const double* mass)
{
    if (!path || !hdr || !io || (hdr->nbody && (!pos || !vel || !mass))) {
        nbh_set_error("nbhip_checkpoint_pack: missing argument");
        return NBHIP_USER_ERROR;
    }
    if (!io->ctx || io->ctx_size == 0 || (io->state_size && !io->state) ||
        (io->n_orbit_trace && !io->orbit_trace) || (io->n_shift_lmc && !io->shift_lmc)) {
        nbh_set_error("nbhip_checkpoint_pack: NBodyCtx image, NBodyState image or vectors missing");
        return NBHIP_USER_ERROR;
    }
    const size_t n = hdr->nbody;
    const size_t hs = header_size(io->ctx_size);
    std::vector<unsigned char> head(hs, 0);
    memcpy(head.data(), kMagic, sizeof(kMagic));
    const uint32_t f[9] = { io->major_version, io->minor_version, hdr->nbody, hdr->step,
                            (uint32_t) sizeof(double), (uint32_t) sizeof(void*),
                            io->n_orbit_trace, io->n_shift_lmc, hdr->tree_incest };
    for (int k = 0; k < 9; ++k) put_u32(head.data() + 128 + 4 * k, f[k]);
    put_f64(head.data() + 168, hdr->rsize);
    memcpy(head.data() + kHdrFixed, io->ctx, io->ctx_size);

    /* the reference writes a temporary file and renames it over the
     * checkpoint (nbWriteCheckpointWithTmpFile, nbody_checkpoint.c:629-662) */
    const std::string tmp = std::string(path) + ".tmp" + std::to_string((long) getpid());
    FILE* fp = fopen(tmp.c_str(), "wb");
    if (!fp) {
        nbh_set_error("nbhip_checkpoint_pack: cannot open '%s': %s", tmp.c_str(), strerror(errno));
        return NBHIP_CHECKPOINT_ERROR;
    }
    bool ok = write_all(fp, head.data(), hs);
    const uint64_t ssz = io->state_size;
    ok = ok && write_all(fp, &ssz, sizeof(ssz)) && write_all(fp, io->state, io->state_size);
    std::vector<unsigned char> rec(kBody);
    for (int tab = 0; tab < 2 && ok; ++tab) {
        const double* xs = tab == 0 ? pos : (io->best_pos ? io->best_pos : pos);
        const double* vs = tab == 0 ? vel : (io->best_vel ? io->best_vel : vel);
        for (size_t i = 0; i < n && ok; ++i) {
            const short type = io->type ? io->type[i] : (short) 1;
            const unsigned id = io->id ? io->id[i] : (unsigned) i;
            put_body(rec.data(), xs + 3 * i, vs + 3 * i, mass[i], type, id);
            ok = write_all(fp, rec.data(), kBody);
        }
    }
    ok = ok && write_all(fp, io->orbit_trace, (size_t) io->n_orbit_trace * kVec);
    if (ok && io->n_shift_lmc) {
        ok = write_all(fp, io->shift_lmc, (size_t) io->n_shift_lmc * kVec) &&
             write_all(fp, io->lmc_pos, kVec) && write_all(fp, io->lmc_vel, kVec);
    }
    ok = ok && write_all(fp, kTail, sizeof(kTail));
    /* pad to the reference's file size (its LMC slot is counted even when empty) */
    const size_t want = file_size(io->ctx_size, io->state_size, n, io->n_orbit_trace,
                                  io->n_shift_lmc);
    if (ok) {
        const long at = ftell(fp);
        if (at < 0 || (size_t) at > want) {
            ok = false;
        } else {
            std::vector<unsigned char> zero(want - (size_t) at, 0);
            ok = write_all(fp, zero.data(), zero.size());
        }
    }
    ok = (fflush(fp) == 0) && ok;
    ok = (fclose(fp) == 0) && ok;
    if (!ok || rename(tmp.c_str(), path) != 0) {
        nbh_set_error("nbhip_checkpoint_pack: writing '%s' failed: %s", path, strerror(errno));
        unlink(tmp.c_str());
        return NBHIP_CHECKPOINT_ERROR;
    }
    return NBHIP_SUCCESS;
}

extern "C" unsigned nbhip_checkpoint_unpack(const char* path, nbhip_checkpoint_header* hdr,
                                            nbhip_checkpoint_io* io, double* pos, double* vel,
                                            double* mass)
{
    if (!path || !hdr || !io || io->ctx_size == 0) {
        nbh_set_error("nbhip_checkpoint_unpack: missing argument");
        return NBHIP_USER_ERROR;
    }
    FILE* fp = fopen(path, "rb");
    if (!fp) {
        nbh_set_error("nbhip_checkpoint_unpack: cannot open '%s': %s", path, strerror(errno));
        return NBHIP_CHECKPOINT_ERROR;
    }
    std::vector<unsigned char> buf;
    if (fseek(fp, 0, SEEK_END) == 0) {
        const long sz = ftell(fp);
        if (sz > 0) {
            buf.resize((size_t) sz);
            rewind(fp);
            if (fread(buf.data(), 1, buf.size(), fp) != buf.size()) buf.clear();
        }
    }
    fclose(fp);
    if (buf.empty()) {
        nbh_set_error("nbhip_checkpoint_unpack: checkpoint '%s' is empty or unreadable", path);
        return NBHIP_CHECKPOINT_ERROR;
    }
    const unsigned char* b = buf.data();
    const size_t hs = header_size(io->ctx_size);
    if (buf.size() < hs + sizeof(uint64_t)) {
        nbh_set_error("nbhip_checkpoint_unpack: '%s' is shorter than its header", path);
        return NBHIP_CHECKPOINT_ERROR;
    }
    if (strncmp((const char*) b, kMagic, 128) != 0) {
        nbh_set_error("nbhip_checkpoint_unpack: no mwnbody header in '%s'", path);
        return NBHIP_CHECKPOINT_ERROR;
    }
    nbhip_checkpoint_header h;
    h.major_version = get_u32(b + 128);
    h.minor_version = get_u32(b + 132);
    h.nbody = get_u32(b + 136);
    h.step = get_u32(b + 140);
    h.real_size = get_u32(b + 144);
    h.ptr_size = get_u32(b + 148);
    h.n_orbit_trace = get_u32(b + 152);
    h.n_shift_lmc = get_u32(b + 156);
    h.tree_incest = get_u32(b + 160);
    h.rsize = get_f64(b + 168);
    uint64_t ssz;
    memcpy(&ssz, b + hs, sizeof(ssz));
    /* nbVerifyCheckpointHeader: size, real size, pointer size, version */
    const size_t want = file_size(io->ctx_size, (size_t) ssz, h.nbody, h.n_orbit_trace,
                                  h.n_shift_lmc);
    if (ssz > buf.size() || want != buf.size()) {
        nbh_set_error("nbhip_checkpoint_unpack: '%s' is %zu bytes, %zu expected for %u bodies",
                      path, buf.size(), want, h.nbody);
        return NBHIP_CHECKPOINT_ERROR;
    }
    if (h.real_size != sizeof(double) || h.ptr_size != sizeof(void*)) {
        nbh_set_error("nbhip_checkpoint_unpack: '%s' has real size %u, pointer size %u", path,
                      h.real_size, h.ptr_size);
        return NBHIP_CHECKPOINT_ERROR;
    }
    if (h.major_version != io->major_version || h.minor_version != io->minor_version) {
        nbh_set_error("nbhip_checkpoint_unpack: '%s' is for version %u.%u, expected %u.%u", path,
                      h.major_version, h.minor_version, io->major_version, io->minor_version);
        return NBHIP_CHECKPOINT_ERROR;
    }
    if (io->state_size && (uint64_t) io->state_size != ssz) {
        nbh_set_error("nbhip_checkpoint_unpack: '%s' holds a %llu-byte NBodyState, %zu expected",
                      path, (unsigned long long) ssz, io->state_size);
        return NBHIP_CHECKPOINT_ERROR;
    }
    size_t off = hs + sizeof(uint64_t) + (size_t) ssz;
    const size_t bodies = off;
    off += 2 * (size_t) h.nbody * kBody;
    const size_t trace = off;
    off += (size_t) h.n_orbit_trace * kVec;
    const size_t shift = off;
    if (h.n_shift_lmc) off += (size_t) h.n_shift_lmc * kVec + 2 * kVec;
    if (off + sizeof(kTail) > buf.size() || memcmp(b + off, kTail, sizeof(kTail)) != 0) {
        nbh_set_error("nbhip_checkpoint_unpack: no end marker in '%s'", path);
        return NBHIP_CHECKPOINT_ERROR;
    }
    *hdr = h;
    if (io->ctx) memcpy(io->ctx, b + kHdrFixed, io->ctx_size);
    if (io->state && io->state_size) memcpy(io->state, b + hs + sizeof(uint64_t), io->state_size);
    for (size_t i = 0; i < h.nbody; ++i) {
        const unsigned char* r = b + bodies + i * kBody;
        double m;
        get_body(r, pos ? pos + 3 * i : nullptr, vel ? vel + 3 * i : nullptr, &m,
                 io->type ? io->type + i : nullptr, io->id ? io->id + i : nullptr);
        if (mass) mass[i] = m;
        const unsigned char* rb = r + (size_t) h.nbody * kBody;
        get_body(rb, io->best_pos ? io->best_pos + 3 * i : nullptr,
                 io->best_vel ? io->best_vel + 3 * i : nullptr, nullptr, nullptr, nullptr);
    }
    /* vectors: copied when the caller's buffers hold them (counts are in hdr) */
    if (io->orbit_trace && io->n_orbit_trace >= h.n_orbit_trace)
        memcpy(io->orbit_trace, b + trace, (size_t) h.n_orbit_trace * kVec);
    if (h.n_shift_lmc) {
        if (io->shift_lmc && io->n_shift_lmc >= h.n_shift_lmc)
            memcpy(io->shift_lmc, b + shift, (size_t) h.n_shift_lmc * kVec);
        memcpy(io->lmc_pos, b + shift + (size_t) h.n_shift_lmc * kVec, kVec);
        memcpy(io->lmc_vel, b + shift + (size_t) h.n_shift_lmc * kVec + kVec, kVec);
    }
    return NBHIP_SUCCESS;
}

extern "C" unsigned nbhip_checkpoint_write(nbhip_state* st, const char* path,
                                           const nbhip_checkpoint_io* io)
{
    if (!st || !path || !io) return NBHIP_USER_ERROR;
    if (st->nsim != 1) {
        nbh_set_error("nbhip_checkpoint_write: a checkpoint file holds one simulation");
        return NBHIP_USER_ERROR;
    }
    const size_t n = (size_t) st->n;
    std::vector<double> x(3 * n), v(3 * n), m(n), bx, bv;
    unsigned rc = nbhip_download(st, x.data(), v.data(), nullptr);
    if (rc) return rc;
    NBH_CHECK(hipSetDevice(st->device));
    if (n) {
        NBH_CHECK(hipMemcpyAsync(m.data(), st->m, n * sizeof(double), hipMemcpyDeviceToHost,
                                 st->stream));
    }
    std::vector<short> type;
    if (!io->type && st->ignore && n) {
        /* Body.type: ignored bodies (the dark matter of a mixed dwarf) are
         * negative, as the histogram's ignoreBody reads them */
        std::vector<unsigned char> ig(n);
        NBH_CHECK(hipMemcpyAsync(ig.data(), st->ignore, n, hipMemcpyDeviceToHost, st->stream));
        NBH_CHECK(hipStreamSynchronize(st->stream));
        type.resize(n);
        for (size_t i = 0; i < n; ++i) type[i] = ig[i] ? (short) -1 : (short) 1;
    }
    NBH_CHECK(hipStreamSynchronize(st->stream));
    nbhip_checkpoint_io w = *io;
    if (!type.empty()) w.type = type.data();
    if (!io->best_pos && !io->best_vel && st->best_kept) {
        bx.resize(3 * n);
        bv.resize(3 * n);
        rc = nbhip_best_bodies(st, bx.data(), bv.data());
        if (rc) return rc;
        w.best_pos = bx.data();
        w.best_vel = bv.data();
    }
    nbhip_tree_info ti;
    rc = nbhip_tree_info_get(st, &ti);
    if (NBHIP_STATUS_IS_FATAL(rc)) return rc;
    nbhip_checkpoint_header h;
    memset(&h, 0, sizeof(h));
    h.nbody = (unsigned) n;
    h.step = st->step;
    h.rsize = ti.rsize;
    h.tree_incest = (st->h_sim_status[0] & NBHIP_TREE_INCEST_NONFATAL) ? 1u : 0u;
    return nbhip_checkpoint_pack(path, &h, &w, x.data(), v.data(), m.data());
}

extern "C" unsigned nbhip_checkpoint_read(nbhip_state* st, const char* path,
                                          nbhip_checkpoint_io* io)
{
    if (!st || !path || !io) return NBHIP_USER_ERROR;
    if (st->nsim != 1) {
        nbh_set_error("nbhip_checkpoint_read: a checkpoint file holds one simulation");
        return NBHIP_USER_ERROR;
    }
    nbhip_checkpoint_header h;
    nbhip_checkpoint_io peek = *io;
    peek.type = nullptr;
    peek.id = nullptr;
    peek.best_pos = peek.best_vel = nullptr;
    peek.orbit_trace = peek.shift_lmc = nullptr;
    unsigned rc = nbhip_checkpoint_unpack(path, &h, &peek, nullptr, nullptr, nullptr);
    if (rc) return rc;
    if ((int) h.nbody != st->n) {
        nbh_set_error("nbhip_checkpoint_read: '%s' holds %u bodies, the state %d", path, h.nbody,
                      st->n);
        return NBHIP_USER_ERROR;
    }
    const size_t n = h.nbody;
    std::vector<double> x(3 * n), v(3 * n), m(n), bx(3 * n), bv(3 * n);
    nbhip_checkpoint_io r = *io;
    r.best_pos = io->best_pos ? io->best_pos : bx.data();
    r.best_vel = io->best_vel ? io->best_vel : bv.data();
    rc = nbhip_checkpoint_unpack(path, &h, &r, x.data(), v.data(), m.data());
    if (rc) return rc;
    memcpy(io->lmc_pos, r.lmc_pos, sizeof(io->lmc_pos));
    memcpy(io->lmc_vel, r.lmc_vel, sizeof(io->lmc_vel));
    rc = nbhip_upload(st, x.data(), v.data(), m.data());
    if (rc) return rc;
    rc = nbhip_restore(st, 0, h.step, h.rsize, h.tree_incest ? NBHIP_TREE_INCEST_NONFATAL : 0u);
    if (rc) return rc;
    /* bestLikelihoodBodyTab back on the device: nbhip_best_bodies returns it */
    NBH_CHECK(hipSetDevice(st->device));
    if (!st->best_buf)
        NBH_CHECK(hipMalloc((void**) &st->best_buf, 6 * (n ? n : 1) * sizeof(double)));
    std::vector<double> soa(6 * n);
    for (size_t i = 0; i < n; ++i)
        for (int d = 0; d < 3; ++d) {
            soa[d * n + i] = r.best_pos[3 * i + d];
            soa[(3 + d) * n + i] = r.best_vel[3 * i + d];
        }
    if (n)
        NBH_CHECK(hipMemcpyAsync(st->best_buf, soa.data(), 6 * n * sizeof(double),
                                 hipMemcpyHostToDevice, st->stream));
    NBH_CHECK(hipStreamSynchronize(st->stream));
    st->best_kept = 1;
    io->n_orbit_trace = h.n_orbit_trace;
    io->n_shift_lmc = h.n_shift_lmc;
    return NBHIP_SUCCESS;
}
