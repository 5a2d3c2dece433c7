// MapGenerator::start() for MI355X: the drop-in for the CUDA render
// generator (src/genmap_cuda_render.cpp:25-200), producing the CPU
// generator's map (src/genmap.cpp:22-175).
//
// This file compiles against the reference's own headers (src/parameters.h,
// src/datareader.h, src/mapgenerator.h) as well as this package's mirrors
// of them: it uses only their public members, and the map constants come
// from dmann_params_of().  Link it with libdmann.so in place of
// genmap_cuda_render.o and chealpixrender.o.
//
// Each DataReader chunk is handed to the library as raw 48-byte records;
// the library gathers chunks in pinned staging buffers and renders them in
// large launches, the copy of one buffer overlapping the render of the
// previous one, so small mem_buffer chunks (2048 in src/VL2_info.ini:57)
// cost no more than large ones.  The per-particle pre-pass that the
// reference runs serially on the host (src/genmap_cuda_render.cpp:102-167)
// runs on the device.  The map stays on the device across chunks and comes
// back once, already divided by dOmega.
//
// Accumulation form: DMANN_PRECISION=fp32 | compensated | fp64 | fixed in
// the environment (default compensated, which meets the 1e-5 per-pixel
// bound against the FP64 CPU generator; fixed gives the same map bit for
// bit on every run).  Flux plug-in: DMANN_FLUX=ann_nosom
// (default) | ann_pwave | ann_somv | ann_somv2 | decay | ann_adiabatic |
// ann_core_nosom (cores from ./halocore/VL2_Halos.data or DMANN_HALO_FILE) -- the
// reference links one src/flux/flux_<kind>.cpp per executable
// (annmap_cuda_nosom, annmap_cuda_pwave, ..., src/Makefile:130-160).
#include <rpc/types.h>
#include <rpc/xdr.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <string>
#include <vector>

#include "dmann.h"
#include "dmann_params_of.h"
#include "mapgenerator.h"

static int precision_from_env()
{
    const char *s = getenv("DMANN_PRECISION");
    if (!s || !*s) return DMANN_PREC_COMPENSATED;
    if (!strcmp(s, "fp32")) return DMANN_PREC_FP32;
    if (!strcmp(s, "fp64")) return DMANN_PREC_FP64;
    if (!strcmp(s, "fixed")) return DMANN_PREC_FIXED;
    return DMANN_PREC_COMPENSATED;
}

static int flux_from_env()
{
    static const char *names[] = {"ann_nosom", "ann_pwave", "ann_somv",      "ann_somv2",
                                  "decay",     "ann_adiabatic", "ann_core_nosom"};
    const char *s = getenv("DMANN_FLUX");
    if (!s || !*s) return DMANN_FLUX_ANN_NOSOM;
    for (int k = 0; k < 7; k++)
        if (!strcmp(s, names[k])) return k;
    printf("Unknown flux plug-in %s\n", s);
    exit(1);
}

void MapGenerator::start()
{
    const int Npix = par_->map.Npix;
    dmann_params p;
    dmann_params_of(*par_, precision_from_env(), &p);
    p.flux_kind = flux_from_env();
    const int64_t cap = par_->memParts > 0 ? par_->memParts : 1;
    dmann_ctx *ctx = nullptr;
    if (dmann_create(&ctx, &p, cap, 0) != DMANN_OK) {
        printf("HIP initialize error! %s\n", dmann_last_error(nullptr));
        exit(1);
    }
    if (p.flux_kind == DMANN_FLUX_ANN_CORE_NOSOM) {
        // the reference's HaloCore reads ./halocore/VL2_Halos.data
        // (src/halocore/halocore.h:16); DMANN_HALO_FILE overrides the path
        const char *hf = getenv("DMANN_HALO_FILE");
        const std::string path = hf && *hf ? hf : "./halocore/VL2_Halos.data";
        const int64_t nh = dmann_read_halo_file(path.c_str(), nullptr, 0);
        if (nh <= 0) {
            printf("Are you input a empty halo file?\n");
            exit(0);
        }
        std::vector<float> halos((size_t)nh * 4);
        dmann_read_halo_file(path.c_str(), halos.data(), nh);
        printf("There are %d halos in consideration!\n", (int)nh);
        if (dmann_set_halo_cores(ctx, halos.data(), nh) != DMANN_OK) {
            printf("HIP initialize error! %s\n", dmann_last_error(ctx));
            exit(1);
        }
    }
    std::cout << "Creating map!!!" << std::endl;
    std::cout << "---10---20---30---40---50---60---70---80---90--100%\n";
    const int Np = reader_->getPartNum();
    const int rec = Np / 50 > 0 ? Np / 50 : 1;
    long long count = 0, marks = 0;
    while (reader_->hasNext()) {
        const int n = reader_->getMemparts();
        if (dmann_render(ctx, reinterpret_cast<const dmann_particle *>(reader_->getBuf()), n) != DMANN_OK) {
            printf("HIP render error! %s\n", dmann_last_error(ctx));
            exit(1);
        }
        count += n;
        for (; marks * rec < count && marks < 50; ++marks) std::cout << "#";
        std::cout.flush();
        reader_->loadBuffer();
    }
    map_ = (MAPTYPE *)calloc((size_t)Npix, sizeof(MAPTYPE));
    const int st = dmann_get_map(ctx, map_);
    if (st != DMANN_OK) {
        printf("HIP map error! %s\n", dmann_last_error(ctx));
        exit(1);
    }
    dmann_stats stats;
    dmann_get_stats(ctx, &stats);
    dmann_destroy(ctx);
    isFinished_ = true;
    std::cout << "\nFinished, " << stats.kept << " particles rendered!." << std::endl;
    double totalC = 0;
    for (int i = 0; i < Npix; i++) totalC += map_[i] * par_->map.dOmega;
    printf("Total Conts: %f\n", totalC);
}
