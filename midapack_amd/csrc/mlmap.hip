// mlmap.hip -- the MLmap entry point of libmappraiser (mappraiser.c:34-380)
// on one rank and one GPU, with its outputs: x2map_pol (mappraiser.c:836-853)
// and the HEALPix FITS maps of write_map (iofiles.c:19-53).  Host code only:
// the solve is mappraiser_gpu_pcg_gls.
//
// The FITS files are written without cfitsio: a primary HDU with no data and
// one BINTABLE extension of 12 nside^2 rows with a single 8-byte float column
// "SIGNAL", the header cards in the order and format cfitsio writes them for
// write_map's calls, big-endian data padded to 2880-byte records.  Apart from
// the DATE value the files are byte-identical to the reference's.
#include "mgpu_internal.h"

#include <chrono>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <string>
#include <unistd.h>
#include <vector>

namespace mgpu {
namespace {

constexpr int kRecord = 2880;

// one 80-character header card; value already formatted (right-justified to
// column 30 for numbers and logicals, quoted strings left-justified from column 11)
void card(std::string &h, const char *key, const std::string &value, const char *comment) {
    char buf[81];
    std::string c = key;
    c.resize(8, ' ');
    c += "= ";
    c += value;
    if (comment) {
        if (c.size() < 30) c.resize(30, ' ');
        c += " / ";
        c += comment;
    }
    c.resize(80, ' ');
    std::memcpy(buf, c.data(), 80);
    h.append(buf, 80);
}

std::string num(long long v) {
    char b[32];
    std::snprintf(b, sizeof(b), "%20lld", v);
    return b;
}

std::string logical(bool v) { return std::string(19, ' ') + (v ? "T" : "F"); }

// FITS string value: quoted, padded with blanks to at least 8 characters
std::string str(const char *v) {
    std::string s = v;
    if (s.size() < 8) s.resize(8, ' ');
    return "'" + s + "'";
}

void comment_card(std::string &h, const char *key, const char *text) {
    std::string c = key;
    c.resize(8, ' ');
    c += text;
    c.resize(80, ' ');
    h += c;
}

void end_header(std::string &h) {
    std::string c = "END";
    c.resize(80, ' ');
    h += c;
    h.resize((h.size() + kRecord - 1) / kRecord * kRecord, ' ');
}

void put_be64(unsigned char *p, double v) {
    uint64_t u;
    std::memcpy(&u, &v, 8);
    for (int i = 0; i < 8; i++) p[i] = (unsigned char) (u >> (56 - 8 * i));
}

}  // namespace
}  // namespace mgpu

using namespace mgpu;

extern "C" {

int mappraiser_gpu_x2map_pol(double *mapI, double *mapQ, double *mapU, double *Cond, int *hits, const double *x,
                             const int *lstid, const double *cond, const int *lhits, int xsize) {
    MGPU_ARG(mapI && mapQ && mapU && Cond && hits && x && lstid && cond && lhits, "NULL argument");
    for (int i = 0; i < xsize; i++) {
        const int p = lstid[i] / 3;
        if (i % 3 == 0) {
            mapI[p] = x[i];
            hits[p] = lhits[i / 3];
            Cond[p] = cond[i / 3];
        } else if (i % 3 == 1) {
            mapQ[p] = x[i];
        } else {
            mapU[p] = x[i];
        }
    }
    return MGPU_OK;
}

int mappraiser_gpu_write_map(const void *signal, int type, long nside, const char *filename, char nest,
                             const char *coordsys) {
    MGPU_ARG(signal && filename && coordsys, "NULL argument");
    MGPU_ARG(type == MGPU_TDOUBLE || type == MGPU_TINT, "type=%d: MGPU_TDOUBLE or MGPU_TINT", type);
    MGPU_ARG(nside > 0 && nside <= 8192, "nside=%ld", nside);
    MGPU_ARG(std::strlen(coordsys) >= 1, "bad coordsys value");
    const long long npix = 12LL * nside * nside;
    // primary HDU: no data (fits_create_img SHORT_IMG, naxis 0, then fits_write_date)
    std::string h;
    card(h, "SIMPLE", logical(true), "file does conform to FITS standard");
    card(h, "BITPIX", num(16), "number of bits per data pixel");
    card(h, "NAXIS", num(0), "number of data axes");
    card(h, "EXTEND", logical(true), "FITS dataset may contain extensions");
    comment_card(h, "COMMENT", "  FITS (Flexible Image Transport System) format is defined in 'Astronomy");
    comment_card(h, "COMMENT", "  and Astrophysics', volume 376, page 359; bibcode: 2001A&A...376..359H");
    char date[32];
    const std::time_t now = std::time(nullptr);
    std::tm tm_utc{};
    gmtime_r(&now, &tm_utc);
    std::strftime(date, sizeof(date), "%Y-%m-%dT%H:%M:%S", &tm_utc);
    card(h, "DATE", str(date), "file creation date (YYYY-MM-DDThh:mm:ss UT)");
    end_header(h);
    // BINTABLE extension (fits_create_tbl, 1 column "SIGNAL" of format 1D) + HEALPix keys
    card(h, "XTENSION", str("BINTABLE"), "binary table extension");
    card(h, "BITPIX", num(8), "8-bit bytes");
    card(h, "NAXIS", num(2), "2-dimensional binary table");
    card(h, "NAXIS1", num(8), "width of table in bytes");
    card(h, "NAXIS2", num(npix), "number of rows in table");
    card(h, "PCOUNT", num(0), "size of special data area");
    card(h, "GCOUNT", num(1), "one data group (required keyword)");
    card(h, "TFIELDS", num(1), "number of fields in each row");
    card(h, "TTYPE1", str("SIGNAL"), "label for field   1");
    card(h, "TFORM1", str("1D"), "data format of field: 8-byte DOUBLE");
    card(h, "TUNIT1", str(" "), "physical unit of field");
    card(h, "EXTNAME", str("BINTABLE"), "name of this binary table extension");
    card(h, "PIXTYPE", str("HEALPIX"), "HEALPIX Pixelisation");
    card(h, "ORDERING", str(nest ? "NESTED  " : "RING    "), "Pixel ordering scheme, either RING or NESTED");
    card(h, "NSIDE", num(nside), "Resolution parameter for HEALPIX");
    const char c0 = coordsys[0];
    const char *cs = c0 == 'G' ? "G       " : c0 == 'E' ? "E       " : "C       ";  // setCoordSysHP
    if (c0 != 'G' && c0 != 'E' && c0 != 'C' && c0 != 'Q')
        std::fprintf(stderr, "write_map: coordinate system '%c' unknown, celestial (C) written\n", c0);
    card(h, "COORDSYS", str(cs), "Pixelisation coordinate system");
    comment_card(h, "COMMENT", "G = Galactic, E = ecliptic, C = celestial = equatorial");
    end_header(h);
    const size_t ndata = (size_t) npix * 8;
    std::vector<unsigned char> data((ndata + kRecord - 1) / kRecord * kRecord, 0);
    for (long long i = 0; i < npix; i++) {
        const double v = type == MGPU_TDOUBLE ? static_cast<const double *>(signal)[i]
                                              : (double) static_cast<const int *>(signal)[i];
        put_be64(&data[(size_t) i * 8], v);
    }
    FILE *fp = std::fopen(filename, "wb");
    MGPU_ARG(fp, "cannot create %s", filename);
    const bool ok = std::fwrite(h.data(), 1, h.size(), fp) == h.size() &&
                    std::fwrite(data.data(), 1, data.size(), fp) == data.size();
    const bool closed = std::fclose(fp) == 0;
    MGPU_ARG(ok && closed, "write error on %s", filename);
    return MGPU_OK;
}

// MLmap / MTmap after their own setup (mappraiser.c:34-380, 387-838): the solve
// of pcg_gls_impl (with template classes when tm is not NULL), the residual
// file, x2map_pol and the five FITS maps
static void map_driver(const char *what, char *outpath, char *ref, int solver, int precond, int Z_2lvl, double tol,
                       int maxiter, int nside, void *data_size_proc, int nb_blocks_loc, void *local_blocks_sizes,
                       int Nnz, void *pix, void *pixweights, void *signal, double *noise, int lambda, double *invtt,
                       const mgpu_templates *tm) {
    std::printf("\n############# MAPPRAISER %s (MI355X drop-in, one rank) ################\n", what);
    if (solver != 0) {
        set_error("solver=%d: only PCG (solver = 0) is available", solver);
        std::fprintf(stderr, "The choice of solver is %d, but only PCG (solver = 0) is available.\n", solver);
        return;
    }
    if (!data_size_proc || !pix || !pixweights || !signal || !noise || !invtt || !local_blocks_sizes) {
        set_error("%s: NULL argument", what);
        std::fprintf(stderr, "%s: NULL argument\n", what);
        return;
    }
    const int m = static_cast<int *>(data_size_proc)[0];
    std::printf("[rank 0] M=%d\n", m);
    std::fflush(stdout);
    const size_t cap = (size_t) m * (Nnz > 0 ? Nnz : 3);
    std::vector<double> x(cap), cond((size_t) m);
    std::vector<int> lhits((size_t) m), lind(cap);
    mgpu_options opt;
    mappraiser_gpu_options_default(&opt);
    opt.precond = precond;
    opt.z_2lvl = Z_2lvl;
    if (const char *e = std::getenv("MGPU_PRECISION"); e && !tm) opt.precision = std::strcmp(e, "mixed") == 0;
    if (const char *e = std::getenv("MGPU_OFFSET_MODE"))
        opt.offset_mode = std::strcmp(e, "symmetric") == 0 ? MGPU_OFFSET_SYMMETRIC
                          : std::strcmp(e, "none") == 0    ? MGPU_OFFSET_NONE
                                                           : MGPU_OFFSET_REFERENCE;
    std::printf("##### Start PCG ####################\n");
    std::fflush(stdout);
    const auto t0 = std::chrono::steady_clock::now();
    mgpu_report rep{};
    const int st = pcg_gls_impl(outpath, ref, m, Nnz, static_cast<int *>(pix), static_cast<double *>(pixweights),
                                static_cast<double *>(signal), noise, nb_blocks_loc,
                                static_cast<int *>(local_blocks_sizes), lambda, invtt, tol, maxiter, x.data(),
                                cond.data(), lhits.data(), lind.data(), &opt, &rep, tm);
    if (st != MGPU_OK) {
        std::fprintf(stderr, "mappraiser_gpu: %s\n", mappraiser_gpu_last_error());
        return;
    }
    const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    std::printf("##### End PCG ####################\n");
    std::printf("[rank 0] Total PCG time=%lf \n", secs);
    std::printf("  -> iterations %d, converged %d, pixels cut %d\n", rep.iterations, rep.converged, rep.n_cut);
    std::fflush(stdout);
    // gather to rank 0 is the identity; x2map_pol into HEALPix maps (mappraiser.c:235-296)
    const int mapsize = rep.n_map;
    const long long npix = 12LL * nside * nside;
    std::vector<int> lstid((size_t) mapsize);
    for (int i = 0; i < mapsize; i++) {
        lstid[i] = lind[(size_t) i + (size_t) Nnz * rep.trash_pix];
        if (lstid[i] < 0 || lstid[i] / 3 >= npix) {
            set_error("map index %d outside the HEALPix map of nside %d", lstid[i], nside);
            std::fprintf(stderr, "mappraiser_gpu: %s\n", mappraiser_gpu_last_error());
            return;
        }
    }
    std::vector<double> mapI((size_t) npix, 0.0), mapQ((size_t) npix, 0.0), mapU((size_t) npix, 0.0),
            Cond((size_t) npix, 0.0);
    std::vector<int> hits((size_t) npix, 0);
    mappraiser_gpu_x2map_pol(mapI.data(), mapQ.data(), mapU.data(), Cond.data(), hits.data(), x.data(),
                             lstid.data(), cond.data(), lhits.data(), mapsize);
    // FITS output (mappraiser.c:297-355): old files are removed first
    std::printf("Checking output directory ... old files will be overwritten\n");
    const char *kinds[5] = {"mapI", "mapQ", "mapU", "Cond", "Hits"};
    std::string names[5];
    int w = 1;
    for (int k = 0; k < 5; k++) {
        names[k] = std::string(outpath ? outpath : ".") + "/" + kinds[k] + "_" + (ref ? ref : "") + ".fits";
        if (access(names[k].c_str(), F_OK) != -1 && std::remove(names[k].c_str()) != 0) {
            std::printf("Error: unable to delete the file %s\n", names[k].c_str());
            w = 0;
        }
    }
    if (w != 1) {
        std::printf("IO Error: Could not overwrite old files, map results will not be stored ;(\n");
        return;
    }
    std::printf("Writing HEALPix maps FITS files to %s...\n", outpath ? outpath : ".");
    const void *maps[5] = {mapI.data(), mapQ.data(), mapU.data(), Cond.data(), hits.data()};
    for (int k = 0; k < 5; k++) {
        if (mappraiser_gpu_write_map(maps[k], k == 4 ? MGPU_TINT : MGPU_TDOUBLE, nside, names[k].c_str(), 1, "C") !=
            MGPU_OK) {
            std::fprintf(stderr, "mappraiser_gpu: %s\n", mappraiser_gpu_last_error());
            return;
        }
    }
    std::fflush(stdout);
}

void MLmap(intptr_t comm, char *outpath, char *ref, int solver, int precond, int Z_2lvl, int pointing_commflag,
           double tol, int maxiter, int enlFac, int ortho_alg, int bs_red, int nside, void *data_size_proc,
           int nb_blocks_loc, void *local_blocks_sizes, int Nnz, void *pix, void *pixweights, void *signal,
           double *noise, int lambda, double *invtt) {
    (void) comm, (void) pointing_commflag, (void) enlFac, (void) ortho_alg, (void) bs_red;
    map_driver("MLmap", outpath, ref, solver, precond, Z_2lvl, tol, maxiter, nside, data_size_proc, nb_blocks_loc,
               local_blocks_sizes, Nnz, pix, pixweights, signal, noise, lambda, invtt, nullptr);
}

void MTmap(intptr_t comm, char *outpath, char *ref, int solver, int precond, int Z_2lvl, int pointing_commflag,
           int npoly, int nhwp, double delta_t, int ground, int n_sss_bins, double tol, int maxiter, int enlFac,
           int ortho_alg, int bs_red, int nside, int **sweeptstamps, int *nsweeps, double **az, double *az_min,
           double *az_max, double **hwp_angle, int nces, void *data_size_proc, int nb_blocks_loc,
           void *local_blocks_sizes, int Nnz, void *pix, void *pixweights, void *signal, double *noise,
           double sampling_freq, double *invtt) {
    (void) comm, (void) pointing_commflag, (void) enlFac, (void) ortho_alg, (void) bs_red;
    // the template classes of mappraiser.c:560-600 (bin_az, the HWP cos/sin and
    // Tlist_init are built inside templates_set) and the white-noise model with
    // lambda_block_avg = 1 (mappraiser.c:531-554)
    mgpu_templates t{};
    t.npoly = npoly;
    t.ground = ground;
    t.n_sss_bins = n_sss_bins;
    t.nhwp = nhwp;
    t.delta_t = delta_t;
    t.sampling_freq = sampling_freq;
    t.nces = nces;
    t.sweeptstamps = sweeptstamps;
    t.nsweeps = nsweeps;
    t.az = az;
    t.az_min = az_min;
    t.az_max = az_max;
    t.hwp_angle = hwp_angle;
    map_driver("MTmap", outpath, ref, solver, precond, Z_2lvl, tol, maxiter, nside, data_size_proc, nb_blocks_loc,
               local_blocks_sizes, Nnz, pix, pixweights, signal, noise, 1, invtt, &t);
}

}  // extern "C"
