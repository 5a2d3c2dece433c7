// MapGenerator::start() for MI355X: the drop-in for the CUDA render
// generator (src/genmap_cuda_render.cpp:25-200), producing the CPU
// generator's map (src/genmap.cpp:22-175).
//
// Each DataReader chunk is handed to the device as raw 48-byte records; the
// per-particle pre-pass that the reference runs serially on the host
// (src/genmap_cuda_render.cpp:102-167) runs on the device.  The map stays on
// the device across chunks and comes back once, already divided by dOmega.
//
// Accumulation form: DMANN_PRECISION=fp32 | compensated | fp64 in the
// environment (default compensated, which meets the 1e-5 per-pixel bound
// against the FP64 CPU generator).
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iostream>

#include "dmann.h"
#include "mapgenerator.h"

static int precision_from_env()
{
    const char *s = getenv("DMANN_PRECISION");
    if (!s || !*s) return DMANN_PREC_COMPENSATED;
    if (!strcmp(s, "fp32")) return DMANN_PREC_FP32;
    if (!strcmp(s, "fp64")) return DMANN_PREC_FP64;
    return DMANN_PREC_COMPENSATED;
}

void MapGenerator::start()
{
    const int Npix = par_->map.Npix;
    dmann_params p;
    par_->toDmannParams(&p, precision_from_env());
    int64_t cap = par_->memParts > 0 ? par_->memParts : 1;
    dmann_ctx *ctx = nullptr;
    if (dmann_create(&ctx, &p, cap, 0) != DMANN_OK) {
        printf("HIP initialize error! %s\n", dmann_last_error(nullptr));
        exit(1);
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
