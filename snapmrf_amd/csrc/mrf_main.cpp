// mrf_main.cpp -- the `mrf` command-line tool: MR fingerprinting dictionary
// generator and image matcher, a drop-in for the reference CLI
// (src/mrf.cu:34-386) running on gfx950 through the snapmrf C ABI.
//
//   mrf [OPTION] <rfpulse.csv> <imgstack.ra>
//
// Same options, defaults, positional-argument rules and RA outputs as the
// reference.  Differences: long options also accept the spellings of the
// reference's help text (--maps, --states, --timepoints, --NATOMS); -G/-B are
// accepted and ignored (launch shapes are chosen per kernel); the image may
// have any number of dimensions >= 2 (voxels = product of dims[1:]); errors
// return an exit status instead of blocking on stdin.  Additions: --precision
// (BF16 / split-BF16 matchers), --stream (generation streamed through the
// FP32 matcher, no dictionary in memory) and --svd FRAC (rank-K compressed
// matching, the MATLAB path's SVD compression).
#include "snapmrf.h"

#include <hip/hip_runtime_api.h>

#include <getopt.h>
#include <sysexits.h>

#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

namespace {

void print_usage()
{
    // src/functions.cu:76-98
    std::fprintf(stderr, "MR fingerprinting dictionary generator and image reconstructor\n");
    std::fprintf(stderr, "Usage: mrf [OPTION] <rfpulse.csv> <imgstack.ra>\n");
    std::fprintf(stderr, "\t-M, --b1map <rafile>\t\t B1 map input RA file\n");
    std::fprintf(stderr, "\t-N, --NATOMS\t\t\t number of atoms\n");
    std::fprintf(stderr, "\t-a, --atoms <rafile>\t\t dictionary atoms output RA file\n");
    std::fprintf(stderr, "\t-m, --maps  <rafile>\t\t parameter map output RA file\n");
    std::fprintf(stderr, "\t-p, --params <rafile>\t\t dictionary parameters\n");
    std::fprintf(stderr, "\t-t, --T1 start:step:end\t\t T1 grid\n");
    std::fprintf(stderr, "\t-s, --T2 start:step:end\t\t T2 grid\n");
    std::fprintf(stderr, "\t-b, --B0 start:step:end\t\t B0 grid\n");
    std::fprintf(stderr, "\t-r, --B1 start:step:end\t\t B1 grid\n");
    std::fprintf(stderr, "\t-w, --states n\t\t\t number of historical states to track\n");
    std::fprintf(stderr, "\t-e, --echo_type n\t\t echo type\n");
    std::fprintf(stderr, "\t-T, --timepoints n\t\t number of time points\n");
    std::fprintf(stderr, "\t-G, --gridsize n\t\t set GPU gridsize (accepted, unused)\n");
    std::fprintf(stderr, "\t-B, --blocksize n\t\t set GPU blocksize (accepted, unused)\n");
    std::fprintf(stderr, "\t    --precision fp32|bf16|bf16_top16|bf16_top32|bf16_top64|split2|split3\t matcher precision (default fp32)\n");
    std::fprintf(stderr, "\t    --stream\t\t\t generate and match block by block (no -a)\n");
    std::fprintf(stderr, "\t    --svd FRAC\t\t\t match in the rank-K SVD space keeping energy fraction FRAC\n");
    std::fprintf(stderr, "\t-h\t\t\t\t show this help\n");
}

int fail(int status, const char *what)
{
    std::fprintf(stderr, "mrf: %s: %s\n", what, mrf_last_error());
    return status;
}

// device copies of the image [nvoxels][nt] and of the B1 map file (optional)
int upload_inputs(const mrf_ra_t &img, size_t nvoxels, size_t nt, const char *in_b1map,
                  mrf_complex **d_sig, float **d_b1)
{
    *d_sig = nullptr;
    *d_b1 = nullptr;
    if (hipMalloc((void **)d_sig, nvoxels * nt * sizeof(mrf_complex)) != hipSuccess ||
        hipMemcpy(*d_sig, img.data, nvoxels * nt * sizeof(mrf_complex),
                  hipMemcpyHostToDevice) != hipSuccess)
        return fail(EX_OSERR, "image upload");
    if (!in_b1map) return 0;
    mrf_ra_t rb;
    if (mrf_ra_read(&rb, in_b1map)) return fail(EX_NOINPUT, "b1map");
    if (rb.size < nvoxels * sizeof(float)) {
        mrf_ra_free(&rb);
        std::fprintf(stderr, "mrf: %s holds fewer than %zu floats\n", in_b1map, nvoxels);
        return EX_DATAERR;
    }
    int rc = 0;
    if (hipMalloc((void **)d_b1, nvoxels * sizeof(float)) != hipSuccess ||
        hipMemcpy(*d_b1, rb.data, nvoxels * sizeof(float), hipMemcpyHostToDevice) != hipSuccess)
        rc = fail(EX_OSERR, "b1map upload");
    mrf_ra_free(&rb);
    return rc;
}

double now()
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

int save_ra(const void *data, uint64_t eltype, uint64_t elbyte, uint64_t d0, uint64_t d1,
            const char *path)
{
    uint64_t dims[2] = {d0, d1};
    mrf_ra_t ra;
    ra.flags = 0;
    ra.eltype = eltype;
    ra.elbyte = elbyte;
    ra.size = d0 * d1 * elbyte;
    ra.ndims = 2;
    ra.dims = dims;
    ra.data = (uint8_t *)data;
    return mrf_ra_write(&ra, path);
}

}  // namespace

int main(int argc, char *argv[])
{
    // defaults, src/mrf.cu:42-63
    const char *in_mrf = "../data/MRF001.csv", *in_img = nullptr, *in_b1map = nullptr;
    const char *out_maps = nullptr, *out_atoms = nullptr, *out_params = nullptr;
    const char *T1 = "20:20:3000", *T2 = "20:10:1000", *B0 = "0", *B1 = "1";
    const char *echo_type = "epg_ssfp";
    const char *precision = "fp32";
    bool stream = false;
    double svd_frac = 0.0;
    int nstates = 101;
    size_t nreps = 1500, NATOMS = 0;

    static const struct option long_options[] = {
        {"b1map", required_argument, nullptr, 'M'},
        {"NATOMS", required_argument, nullptr, 'N'},
        {"atoms", required_argument, nullptr, 'a'},
        {"map", required_argument, nullptr, 'm'},
        {"maps", required_argument, nullptr, 'm'},
        {"params", required_argument, nullptr, 'p'},
        {"T1", required_argument, nullptr, 't'},
        {"T2", required_argument, nullptr, 's'},
        {"B0", required_argument, nullptr, 'b'},
        {"B1", required_argument, nullptr, 'r'},
        {"nstates", required_argument, nullptr, 'w'},
        {"states", required_argument, nullptr, 'w'},
        {"echo_type", required_argument, nullptr, 'e'},
        {"nreps", required_argument, nullptr, 'T'},
        {"timepoints", required_argument, nullptr, 'T'},
        {"gridsize", required_argument, nullptr, 'G'},
        {"blocksize", required_argument, nullptr, 'B'},
        {"precision", required_argument, nullptr, 'P'},
        {"stream", no_argument, nullptr, 'S'},
        {"svd", required_argument, nullptr, 'V'},
        {"help", no_argument, nullptr, 'h'},
        {nullptr, 0, nullptr, 0}};
    opterr = 0;
    int c, option_index = 0;
    while ((c = getopt_long(argc, argv, "M:N:a:m:p:t:s:b:r:w:e:T:G:B:h", long_options,
                            &option_index)) != -1) {
        switch (c) {
        case 'M': in_b1map = optarg; break;
        case 'N': NATOMS = (size_t)std::atoi(optarg); break;
        case 'a': out_atoms = optarg; break;
        case 'm': out_maps = optarg; break;
        case 'p': out_params = optarg; break;
        case 't': T1 = optarg; break;
        case 's': T2 = optarg; break;
        case 'b': B0 = optarg; break;
        case 'r': B1 = optarg; break;
        case 'w': nstates = std::atoi(optarg); break;
        case 'e': echo_type = optarg; break;
        case 'T': nreps = (size_t)std::atoi(optarg); break;
        case 'G':
        case 'B': break;
        case 'P': precision = optarg; break;
        case 'S': stream = true; break;
        case 'V': svd_frac = std::atof(optarg); break;
        case 'h':
        default: print_usage(); return 1;
        }
    }
    // positional arguments: first ".ra" is the image, first ".csv" the
    // sequence (src/mrf.cu:143-166)
    int csv_count = 0, ra_count = 0;
    for (int i = optind; i < argc; i++) {
        const char *a = argv[i];
        if (std::strstr(a, ".ra") && ra_count == 0) { in_img = a; ra_count++; }
        if (std::strstr(a, ".csv") && csv_count == 0) { in_mrf = a; csv_count++; }
        if (csv_count + ra_count == 2) break;
    }
    int mode = MRF_MATCH_FP32;
    if (!std::strcmp(precision, "bf16")) mode = MRF_MATCH_BF16_TOP8;
    else if (!std::strcmp(precision, "bf16_top16")) mode = MRF_MATCH_BF16_TOP16;
    else if (!std::strcmp(precision, "bf16_top32")) mode = MRF_MATCH_BF16_TOP32;
    else if (!std::strcmp(precision, "bf16_top64")) mode = MRF_MATCH_BF16_TOP64;
    else if (!std::strcmp(precision, "split2")) mode = MRF_MATCH_SPLIT2;
    else if (!std::strcmp(precision, "split3")) mode = MRF_MATCH_SPLIT3;
    else if (std::strcmp(precision, "fp32")) {
        std::fprintf(stderr, "mrf: unknown precision '%s'\n", precision);
        return EX_USAGE;
    }

    if (svd_frac > 0.0 && mode != MRF_MATCH_FP32) {
        std::fprintf(stderr, "mrf: --svd uses the FP32 matcher\n");
        return EX_USAGE;
    }
    if (stream && (svd_frac > 0.0 || out_atoms || !in_img)) {
        std::fprintf(stderr, "mrf: --stream needs an image and excludes -a and --svd\n");
        return EX_USAGE;
    }
    if (svd_frac < 0.0 || svd_frac > 1.0) {
        std::fprintf(stderr, "mrf: --svd takes an energy fraction in (0, 1]\n");
        return EX_USAGE;
    }

    std::printf("\nnstates : %d\nEcho : %s\n\n", nstates, echo_type);

    // grids, src/mrf.cu:175-226
    size_t l_t1, l_t2, l_b0, l_b1;
    std::vector<float> t1, t2, b0, b1;
    if (NATOMS != 0) {
        std::printf("Hard coding for parameters...\n");
        mrf_grid_natoms_mode(NATOMS, &l_t1, &l_t2, &l_b0, &l_b1, nullptr, nullptr, nullptr,
                             nullptr);
        t1.resize(l_t1 + 1);
        t2.resize(l_t2 + 1);
        b0.resize(l_b0 + 1);
        b1.resize(l_b1 + 1);
        mrf_grid_natoms_mode(NATOMS, &l_t1, &l_t2, &l_b0, &l_b1, t1.data(), t2.data(),
                             b0.data(), b1.data());
    } else {
        std::printf("Parsing parameters...\nT1 : %s\nT2 : %s\nB0 : %s\nB1 : %s\n", T1, T2, B0, B1);
        const char *g[4] = {T1, T2, B0, B1};
        size_t *lens[4] = {&l_t1, &l_t2, &l_b0, &l_b1};
        std::vector<float> *vals[4] = {&t1, &t2, &b0, &b1};
        for (int k = 0; k < 4; k++) {
            *lens[k] = mrf_parse_length(g[k]);
            if (*lens[k] == 0) return fail(EX_USAGE, "grid");
            vals[k]->resize(*lens[k]);
            if (mrf_parse_params(vals[k]->data(), g[k], *lens[k])) return fail(EX_USAGE, "grid");
        }
    }
    NATOMS = mrf_compute_natoms(t1.data(), t2.data(), l_t1, l_t2, l_b0, l_b1);
    const size_t natoms = l_b1 ? NATOMS / l_b1 : 0;
    std::printf("l_t1: %zu, l_t2: %zu, l_b0: %zu, l_b1: %zu\n", l_t1, l_t2, l_b0, l_b1);
    std::printf("NATOMS: %zu, natoms: %zu\n", NATOMS, natoms);
    std::printf("Removing situations when T1 <= T2\nNATOMS might be smaller than input\n");
    if (NATOMS == 0) {
        std::fprintf(stderr, "mrf: Number of atoms is 0!\n");
        return EX_SOFTWARE;
    }
    std::vector<float> params(NATOMS * 4);
    if (mrf_trans_params(params.data(), t1.data(), t2.data(), b0.data(), b1.data(), l_t1, l_t2,
                         l_b0, l_b1, NATOMS, 4))
        return fail(EX_SOFTWARE, "trans_params");

    // sequence, src/mrf.cu:258-274
    std::printf("\nInput mrf: %s\n", in_mrf);
    const size_t dim = mrf_csv_dim(in_mrf);
    if (dim == (size_t)-1 || dim < 2) return fail(EX_NOINPUT, "csv");
    const size_t NREPS = dim - 1;
    if (nreps > NREPS) nreps = NREPS;
    std::printf("NREPS: %zu, nreps: %zu\n", NREPS, nreps);
    std::vector<float> h_mrf(NREPS * 4);
    int crc = mrf_csv_read(h_mrf.data(), NREPS, in_mrf);
    if (crc < 0) return fail(EX_DATAERR, "csv");
    if (crc > 0) std::fprintf(stderr, "mrf: warning: %s\n", mrf_last_error());

    float *d_params = nullptr;
    mrf_complex *d_atoms = nullptr;
    if (hipMalloc((void **)&d_params, NATOMS * 4 * sizeof(float)) != hipSuccess ||
        (!stream &&
         hipMalloc((void **)&d_atoms, nreps * NATOMS * sizeof(mrf_complex)) != hipSuccess)) {
        std::fprintf(stderr, "mrf: device allocation failed\n");
        return EX_OSERR;
    }
    if (hipMemcpy(d_params, params.data(), NATOMS * 4 * sizeof(float), hipMemcpyHostToDevice) !=
        hipSuccess)
        return fail(EX_OSERR, "parameter upload");

    double t0 = now();
    if (!stream) {
        std::printf("Constructing dictionary...\n");
        if (int rc = mrf_dict(d_atoms, d_params, h_mrf.data(), nreps, NREPS, NATOMS, natoms, 4,
                              nstates, echo_type))
            return fail(rc == MRF_E_ECHO ? EX_USAGE : EX_SOFTWARE, "dictionary");
        std::printf("Elapsed dictionary time: %.2f s\n\n", now() - t0);
    }

    if (!in_img && out_maps) std::printf("NO data specified! Abort matching!\n");
    if (in_img) {
        std::printf("Input img: %s\nReading %s\n", in_img, in_img);
        mrf_ra_t img;
        if (mrf_ra_read(&img, in_img)) return fail(EX_NOINPUT, "image");
        if (img.ndims < 2 || img.eltype != 4 || img.elbyte != 8) {
            std::fprintf(stderr, "mrf: %s must be a complex64 RA file with >= 2 dims\n", in_img);
            return EX_DATAERR;
        }
        const size_t nt = img.dims[0];
        size_t nvoxels = 1;
        for (uint64_t d = 1; d < img.ndims; d++) nvoxels *= img.dims[d];
        std::printf("nt: %zu, nvoxels: %zu\n\n", nt, nvoxels);
        float *d_maps = nullptr;
        if (hipMalloc((void **)&d_maps, 5 * nvoxels * sizeof(float)) != hipSuccess) {
            std::fprintf(stderr, "mrf: device allocation failed\n");
            return EX_OSERR;
        }
        std::printf("Matching...\n");
        t0 = now();
        if (mode == MRF_MATCH_FP32 && !stream && svd_frac <= 0.0) {
            if (mrf_match(d_maps, (const mrf_complex *)img.data, d_atoms, d_params, nreps, nt,
                          NATOMS, natoms, nvoxels, 4, 5, in_b1map))
                return fail(EX_SOFTWARE, "matching");
        } else {
            // extended entry points: device-resident signals and B1 map
            if (nt != nreps) {
                std::fprintf(stderr, "mrf: image frames (%zu) must equal nreps (%zu)\n", nt, nreps);
                return EX_DATAERR;
            }
            mrf_complex *d_sig = nullptr;
            float *d_b1 = nullptr;
            if (int rc = upload_inputs(img, nvoxels, nt, in_b1map, &d_sig, &d_b1)) return rc;
            if (stream) {
                std::printf("Constructing and matching the dictionary block by block...\n");
                if (int rc = mrf_dict_match_ex(nullptr, echo_type, h_mrf.data(), nreps, NREPS,
                                               d_params, NATOMS, natoms, nstates, nullptr, mode,
                                               d_sig, nvoxels, d_b1, 0, d_maps, nullptr))
                    return fail(rc == MRF_E_ECHO ? EX_USAGE : EX_SOFTWARE, "streamed matching");
            } else if (svd_frac > 0.0) {
                // per-slice bases when the B1 map picks each voxel's slice
                const int per_slice = (d_b1 && NATOMS / natoms > 1) ? 1 : 0;
                const size_t nb = per_slice ? NATOMS / natoms : 1;
                const size_t kmax = nreps < 256 ? nreps : 256;
                mrf_complex *basis = nullptr, *catoms = nullptr;
                size_t K = 0;
                if (hipMalloc((void **)&basis, nb * kmax * nreps * sizeof(mrf_complex)) != hipSuccess)
                    return fail(EX_OSERR, "device allocation");
                if (mrf_svd_basis(nullptr, d_atoms, nreps, NATOMS, natoms, per_slice, svd_frac,
                                  kmax, basis, nullptr, &K))
                    return fail(EX_SOFTWARE, "SVD basis");
                std::printf("SVD rank: %zu (%s)\n", K, per_slice ? "per B1 slice" : "one basis");
                if (hipMalloc((void **)&catoms, NATOMS * K * sizeof(mrf_complex)) != hipSuccess)
                    return fail(EX_OSERR, "device allocation");
                if (mrf_svd_compress(nullptr, basis, K, nreps, per_slice, d_atoms, NATOMS, natoms,
                                     catoms) ||
                    mrf_svd_match_ex(nullptr, basis, K, nreps, per_slice, catoms, d_params, NATOMS,
                                     natoms, d_sig, nvoxels, d_b1, d_maps, nullptr))
                    return fail(EX_SOFTWARE, "SVD matching");
                (void)hipDeviceSynchronize();
                (void)hipFree(basis);
                (void)hipFree(catoms);
            } else {
                void *hi = nullptr, *lo = nullptr;
                const size_t lowp_bytes = NATOMS * mrf_lowp_ld(nreps) * 4;
                const bool need_lo = mode == MRF_MATCH_SPLIT3 || mode == MRF_MATCH_SPLIT2;
                if (hipMalloc(&hi, lowp_bytes) != hipSuccess ||
                    (need_lo && hipMalloc(&lo, lowp_bytes) != hipSuccess))
                    return fail(EX_OSERR, "device allocation");
                if (mrf_lowp_prepare(nullptr, d_atoms, nreps, NATOMS, hi, lo) ||
                    mrf_match_ex(nullptr, mode, d_atoms, hi, lo, d_params, nreps, NATOMS, natoms,
                                 d_sig, nvoxels, d_b1, d_maps, nullptr))
                    return fail(EX_SOFTWARE, "matching");
                (void)hipDeviceSynchronize();
                (void)hipFree(hi);
                if (lo) (void)hipFree(lo);
            }
            if (hipDeviceSynchronize() != hipSuccess) return fail(EX_SOFTWARE, "matching");
            (void)hipFree(d_sig);
            if (d_b1) (void)hipFree(d_b1);
        }
        std::printf("Elapsed matching time: %.2f s\n\n", now() - t0);
        if (!out_maps) out_maps = "../result/maps.ra";
        std::printf("Output maps: %s\n", out_maps);
        std::vector<float> h_map(5 * nvoxels);
        if (hipMemcpy(h_map.data(), d_maps, 5 * nvoxels * sizeof(float),
                      hipMemcpyDeviceToHost) != hipSuccess)
            return fail(EX_OSERR, "maps download");
        if (save_ra(h_map.data(), 3, 4, 5, nvoxels, out_maps)) return fail(EX_CANTCREAT, "maps");
        (void)hipFree(d_maps);
        mrf_ra_free(&img);
    }
    if (out_atoms) {
        std::printf("Output atoms: %s\n", out_atoms);
        std::vector<mrf_complex> h_atoms(nreps * NATOMS);
        if (hipMemcpy(h_atoms.data(), d_atoms, nreps * NATOMS * sizeof(mrf_complex),
                      hipMemcpyDeviceToHost) != hipSuccess)
            return fail(EX_OSERR, "atoms download");
        if (save_ra(h_atoms.data(), 4, 8, nreps, NATOMS, out_atoms)) return fail(EX_CANTCREAT, "atoms");
    }
    if (out_params) {
        std::printf("Output params: %s\n", out_params);
        if (save_ra(params.data(), 3, 4, 4, NATOMS, out_params)) return fail(EX_CANTCREAT, "params");
    }
    if (d_atoms) (void)hipFree(d_atoms);
    (void)hipFree(d_params);
    return 0;
}
