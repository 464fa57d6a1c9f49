// bh_cont.cpp -- the hydro code's continuation files, read and written on
// the host: fort.12.<tstep>.<rank>, one per MPI rank, holding that rank's
// blocks of s, t, a, rho, tau, pot, tempa, frac1, frac2 as one unformatted
// sequential record (ritecont_frac_new.F90:140-152; read back by
// setup_frac.F90:625-641), and fort.13.<tstep>, the root rank's scalars
// (ritecont_frac_new.F90:154-158, setup_frac.F90:645-653).
//
// A rank's block is (numr_dd, numz_dd, numphi), Fortran order, ghost layers
// included: numr_dd = (numr - 2) / numr_procs + 2, and rank = column +
// numr_procs * row owns global J = column (numr_dd - 2) + j, K = row
// (numz_dd - 2) + k (main.F90:148-244).  A record is the payload between two
// 4-byte little-endian markers that hold its length in bytes.
#include "bh_plan_impl.h"

#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

using namespace bh;

namespace {

struct Blocks {
    int numr, numz, numphi, nrp, nzp, nrd, nzd;
    size_t block() const { return (size_t)nrd * nzd * numphi; }
};

int check_layout(const bh_cont_layout *lay, Blocks *b)
{
    if (!lay) return fail(BH_E_INVALID, "null layout");
    if (lay->numr < 3 || lay->numz < 3 || lay->numphi < 1 || lay->numr_procs < 1 || lay->numz_procs < 1)
        return fail(BH_E_INVALID, "continuation layout: bad sizes");
    if ((lay->numr - 2) % lay->numr_procs || (lay->numz - 2) % lay->numz_procs)
        return fail(BH_E_INVALID, "continuation layout: numr - 2 and numz - 2 must divide by the process grid");
    b->numr = lay->numr;
    b->numz = lay->numz;
    b->numphi = lay->numphi;
    b->nrp = lay->numr_procs;
    b->nzp = lay->numz_procs;
    b->nrd = (lay->numr - 2) / lay->numr_procs + 2;
    b->nzd = (lay->numz - 2) / lay->numz_procs + 2;
    if (b->block() * BH_CONT_NFIELDS * sizeof(double) > 0x7fffffffu)
        return fail(BH_E_INVALID, "continuation layout: a rank's record would exceed 2 GiB");
    return BH_OK;
}

// fort.12.<tstep>. as ritecont_frac_new.F90:119-131 builds it: the step in
// i5 (right-justified, blank-padded), i6 or i7, then a dot
std::string step_prefix(const char *dir, const char *stem, int tstep)
{
    char num[32];
    if (tstep < 100000) std::snprintf(num, sizeof num, "%5d", tstep);
    else if (tstep < 1000000) std::snprintf(num, sizeof num, "%6d", tstep);
    else std::snprintf(num, sizeof num, "%7d", tstep);
    std::string d = dir ? dir : "";
    if (!d.empty() && d.back() != '/') d += '/';
    return d + stem + num;
}

bool put_marker(FILE *f, size_t n)
{
    const unsigned char m[4] = {(unsigned char)(n & 0xff), (unsigned char)((n >> 8) & 0xff),
                                (unsigned char)((n >> 16) & 0xff), (unsigned char)((n >> 24) & 0xff)};
    return std::fwrite(m, 1, 4, f) == 4;
}

bool get_marker(FILE *f, size_t *n)
{
    unsigned char m[4];
    if (std::fread(m, 1, 4, f) != 4) return false;
    *n = (size_t)m[0] | ((size_t)m[1] << 8) | ((size_t)m[2] << 16) | ((size_t)m[3] << 24);
    return true;
}

}  // namespace

extern "C" int bh_cont_write(const char *dir, int tstep, const bh_cont_layout *lay, const double *const *fields)
{
    Blocks b;
    int rc = check_layout(lay, &b);
    if (rc != BH_OK) return rc;
    if (!fields) return fail(BH_E_INVALID, "null fields");
    const std::string pre = step_prefix(dir, "fort.12.", tstep) + ".";
    std::vector<double> blk(b.block());
    const size_t payload = b.block() * BH_CONT_NFIELDS * sizeof(double);
    for (int rank = 0; rank < b.nrp * b.nzp; ++rank) {
        const int col = rank % b.nrp, row = rank / b.nrp;
        const std::string path = pre + std::to_string(rank);
        FILE *f = std::fopen(path.c_str(), "wb");
        if (!f) return fail(BH_E_INVALID, "cannot create " + path);
        bool ok = put_marker(f, payload);
        for (int q = 0; q < BH_CONT_NFIELDS && ok; ++q) {
            const double *g = fields[q];
            for (int l = 0; l < b.numphi; ++l)
                for (int k = 0; k < b.nzd; ++k)
                    for (int j = 0; j < b.nrd; ++j) {
                        const int J = col * (b.nrd - 2) + j, K = row * (b.nzd - 2) + k;  // 0-based global
                        blk[((size_t)l * b.nzd + k) * b.nrd + j] =
                            g ? g[((size_t)l * b.numz + K) * b.numr + J] : 0.0;
                    }
            ok = std::fwrite(blk.data(), sizeof(double), blk.size(), f) == blk.size();
        }
        ok = ok && put_marker(f, payload);
        ok = (std::fclose(f) == 0) && ok;
        if (!ok) return fail(BH_E_INVALID, "write failed: " + path);
    }
    return BH_OK;
}

extern "C" int bh_cont_read(const char *prefix, const bh_cont_layout *lay, double *const *fields)
{
    Blocks b;
    int rc = check_layout(lay, &b);
    if (rc != BH_OK) return rc;
    if (!prefix || !fields) return fail(BH_E_INVALID, "null argument");
    std::vector<double> blk(b.block());
    const size_t payload = b.block() * BH_CONT_NFIELDS * sizeof(double);
    for (int rank = 0; rank < b.nrp * b.nzp; ++rank) {
        const int col = rank % b.nrp, row = rank / b.nrp;
        const std::string path = std::string(prefix) + std::to_string(rank);
        FILE *f = std::fopen(path.c_str(), "rb");
        if (!f) return fail(BH_E_INVALID, "cannot open " + path);
        size_t m0 = 0, m1 = 0;
        bool ok = get_marker(f, &m0) && m0 == payload;
        if (!ok) {
            std::fclose(f);
            return fail(BH_E_INVALID, path + ": record of " + std::to_string(m0) + " bytes, the layout needs " +
                                          std::to_string(payload));
        }
        // the rank's interior, and the global boundary layers (its ghosts)
        // on the ranks at the edges of the process grid
        const int jlo = col == 0 ? 0 : 1, jhi = col == b.nrp - 1 ? b.nrd - 1 : b.nrd - 2;
        const int klo = row == 0 ? 0 : 1, khi = row == b.nzp - 1 ? b.nzd - 1 : b.nzd - 2;
        for (int q = 0; q < BH_CONT_NFIELDS && ok; ++q) {
            ok = std::fread(blk.data(), sizeof(double), blk.size(), f) == blk.size();
            double *g = fields[q];
            if (!ok || !g) continue;
            for (int l = 0; l < b.numphi; ++l)
                for (int k = klo; k <= khi; ++k)
                    for (int j = jlo; j <= jhi; ++j) {
                        const int J = col * (b.nrd - 2) + j, K = row * (b.nzd - 2) + k;
                        g[((size_t)l * b.numz + K) * b.numr + J] = blk[((size_t)l * b.nzd + k) * b.nrd + j];
                    }
        }
        ok = ok && get_marker(f, &m1) && m1 == payload;
        std::fclose(f);
        if (!ok) return fail(BH_E_INVALID, path + ": truncated or malformed record");
    }
    return BH_OK;
}

namespace {
// fort.13's record: time, frnum (a default integer, 4 bytes), phi_com, R_com,
// com(3,3), v_com(3), a_com(3), cylin_a_com(3,numphi), cylin_v_com(3,numphi),
// delt(3) (ritecont_frac_new.F90:157-158)
size_t scalars_bytes(int numphi) { return 8 + 4 + 8 + 8 + 72 + 24 + 24 + 2 * 24 * (size_t)numphi + 24; }
}  // namespace

extern "C" int bh_cont_write_scalars(const char *dir, int tstep, int numphi, const bh_cont_scalars *sc)
{
    if (!sc || !sc->cylin_a_com || !sc->cylin_v_com || numphi < 1) return fail(BH_E_INVALID, "bad argument");
    const std::string path = step_prefix(dir, "fort.13.", tstep);
    std::vector<unsigned char> rec;
    auto put = [&](const void *p, size_t n) {
        const unsigned char *c = (const unsigned char *)p;
        rec.insert(rec.end(), c, c + n);
    };
    const int frnum = sc->frnum;
    put(&sc->time, 8);
    put(&frnum, 4);
    put(&sc->phi_com, 8);
    put(&sc->R_com, 8);
    put(sc->com, 72);
    put(sc->v_com, 24);
    put(sc->a_com, 24);
    put(sc->cylin_a_com, 24 * (size_t)numphi);
    put(sc->cylin_v_com, 24 * (size_t)numphi);
    put(sc->delt, 24);
    FILE *f = std::fopen(path.c_str(), "wb");
    if (!f) return fail(BH_E_INVALID, "cannot create " + path);
    bool ok = put_marker(f, rec.size()) && std::fwrite(rec.data(), 1, rec.size(), f) == rec.size() &&
              put_marker(f, rec.size());
    ok = (std::fclose(f) == 0) && ok;
    return ok ? BH_OK : fail(BH_E_INVALID, "write failed: " + path);
}

extern "C" int bh_cont_read_scalars(const char *path, int numphi, bh_cont_scalars *sc)
{
    if (!path || !sc || !sc->cylin_a_com || !sc->cylin_v_com || numphi < 1) return fail(BH_E_INVALID, "bad argument");
    FILE *f = std::fopen(path, "rb");
    if (!f) return fail(BH_E_INVALID, std::string("cannot open ") + path);
    const size_t n = scalars_bytes(numphi);
    std::vector<unsigned char> rec(n);
    size_t m0 = 0, m1 = 0;
    bool ok = get_marker(f, &m0) && m0 == n && std::fread(rec.data(), 1, n, f) == n && get_marker(f, &m1) && m1 == n;
    std::fclose(f);
    if (!ok) return fail(BH_E_INVALID, std::string(path) + ": not a fort.13 record for numphi = " + std::to_string(numphi));
    size_t o = 0;
    auto get = [&](void *p, size_t k) {
        std::memcpy(p, rec.data() + o, k);
        o += k;
    };
    int frnum = 0;
    get(&sc->time, 8);
    get(&frnum, 4);
    sc->frnum = frnum;
    get(&sc->phi_com, 8);
    get(&sc->R_com, 8);
    get(sc->com, 72);
    get(sc->v_com, 24);
    get(sc->a_com, 24);
    get(sc->cylin_a_com, 24 * (size_t)numphi);
    get(sc->cylin_v_com, 24 * (size_t)numphi);
    get(sc->delt, 24);
    return BH_OK;
}

extern "C" int bh_set_potential_from_cont(bh_plan *plan, const char *prefix, int numr_procs, int numz_procs)
{
    if (!plan || !prefix) return fail(BH_E_INVALID, "null argument");
    const Geo &g = plan->g;
    bh_cont_layout lay{g.numr, g.numz, g.numphi, numr_procs, numz_procs};
    std::vector<double> pot((size_t)g.numr * g.numz * g.numphi);
    double *fields[BH_CONT_NFIELDS] = {};
    fields[5] = pot.data();  // pot: the sixth array of the record
    int rc = bh_cont_read(prefix, &lay, fields);
    if (rc != BH_OK) return rc;
    return bh_set_potential(plan, pot.data(), 0);
}
