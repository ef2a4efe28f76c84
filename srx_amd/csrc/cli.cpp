// cli.cpp -- `srx <deck>`: the drop-in for the reference's `trial <deck>` (evp.c:7-31).
// Runs every increment of the deck on the GPU and writes vm.out, dd.out, err.out,
// drx.out and the .iout field dumps of the deck's PrintControl into the current
// directory, as the reference does.
#include "../../include/srx.h"

#include <cstdio>
#include <cstdlib>
#include <cstring>

int main(int argc, char **argv) {
    if (argc < 2) {
        std::fprintf(stderr, "usage: %s <deck> [--mixed] [--no-fields] [--verbose] [--gnd]\n", argv[0]);
        return 1;
    }
    srx_options o;
    srx_default_options(&o);
    o.write_files = 1;
    o.write_fields = 1;
    for (int i = 2; i < argc; ++i) {
        if (!std::strcmp(argv[i], "--mixed")) o.precision = SRX_PREC_MIXED;
        if (!std::strcmp(argv[i], "--no-fields")) o.write_fields = 0;
        if (!std::strcmp(argv[i], "--verbose")) o.verbose = 1;
        if (!std::strcmp(argv[i], "--gnd")) o.gnd = 1;   // the GND path of a DD_GND build
    }
    srx_ctx *ctx = nullptr;
    int rc = srx_open(argv[1], ".", &o, &ctx);
    if (rc) {
        std::fprintf(stderr, "srx: %s (code %d)\n", srx_strerror(rc), rc);
        return rc > 255 ? 1 : rc;
    }
    std::printf("\n\n======================================\n-------------Simulation starts----------------\n");
    srx_step_report r;
    for (;;) {
        rc = srx_step(ctx, &r);
        if (rc == SRX_E_DONE || rc == SRX_E_CREEP_STOP) { rc = 0; break; }
        if (rc) break;
        std::printf("STEP = %d  iterations = %d  Err_s = %e  Err_e = %e  s_vm = %e\n", r.istep, r.iters, r.err_s,
                    r.err_e, r.s_vm);
    }
    std::printf("-------------Simulation ends----------------\n======================================\n\n");
    if (rc) std::fprintf(stderr, "srx: %s (code %d)\n", srx_strerror(rc), rc);
    srx_close(ctx);
    return rc ? 1 : 0;
}
