/*
 * The drop-in header as libdvbrx's callers include it: dvbrx/ldpc.h, then
 * dvbrx/fec.h (rx.c and rxmt.c reach drx_fec_lookup and drx_bch_* through
 * ldpc.h).  Built twice by the tests: with libdvbrx's include directory
 * after this repository's (the reference's fec.h is the one used), and
 * with this repository's alone (the header declares the FEC functions
 * itself, libldpcgpu.so defines them).  Prints the parameters of each rate
 * given on the command line.
 */
#include "dvbrx/ldpc.h"
#if defined(WITH_FEC_H)
#include "dvbrx/fec.h"
#endif

#include <stdio.h>
#include <stdlib.h>

int main(int argc, char** argv)
{
    for (int a = 1; a + 1 < argc; a += 2) {
        drx_fec_params p;
        const int rc = drx_fec_lookup(atoi(argv[a]), argv[a + 1], &p);
        printf("%s %s %d %d %d %d %d\n", argv[a], argv[a + 1], rc, rc ? 0 : p.nldpc, rc ? 0 : p.kldpc,
               rc ? 0 : p.kbch, rc ? 0 : p.t);
    }
    return 0;
}
