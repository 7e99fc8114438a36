/* Test infrastructure: the reference's DVB-T2 FEC job (libdvbrx/src/dvbt2.c
 * decode_job, :2188-2400, with setup_plp_fec, :1953-2082) made callable for
 * the cross-checks.  Both functions are static in that file, so this unit
 * compiles the file where it lies (-I on the reference's src/) and adds one
 * exported entry point; nothing of the reference is copied here.  Built
 * into oracle/_ref/ by oracle/Makefile's `ref` target. */
#include "dvbt2.c"

/* One PLP's FEC blocks, as the reference's FEC thread decodes them:
 * cells (complex float, nblocks * N / m of them, time-deinterleaved) and
 * their channel state -> BBFRAMEs (kbch / 8 bytes a block), BCH success and
 * the MER terms of each block.  With llr != NULL and nblocks at most the
 * reference's SIMD lane count, llr receives each block's int8 soft values in
 * codeword order (the decoder input).  Returns 0, or -1 for an unsupported
 * PLP. */
int t2ref_decode(int cod, int mod, int fec_type, int rotation, int til_length, int til_type, int nblocks,
                 const float* cells, const float* csi, int max_iter, uint8_t* bb, uint8_t* ok, float* err,
                 float* sig, int8_t* llr)
{
    struct drx_dvbt2* dd = calloc(1, sizeof *dd);
    struct fecwork* w = calloc(1, sizeof *w);
    struct fecjob job;
    int rc = -1;
    memset(&job, 0, sizeof job);
    if (!dd || !w)
        goto out;
    w->d = dd;
    w->bitmap_key = -1;
    for (int i = 0; i < 64; i++) {
        w->llr[i] = malloc(64800);
        w->cw[i] = malloc(64800 / 8);
        if (!w->llr[i] || !w->cw[i])
            goto out;
    }
    job.q.cod = cod;
    job.q.mod = mod;
    job.q.fec_type = fec_type;
    job.q.rotation = rotation;
    job.q.time_il_length = til_length;
    job.q.time_il_type = til_type;
    job.q.num_blocks = nblocks;
    job.buf = (cf*)cells;
    job.csi = (float*)csi;
    decode_job(w, &job, max_iter);
    if (job.nblocks != nblocks)
        goto out;
    memcpy(bb, job.bb, (size_t)nblocks * job.kb);
    memcpy(ok, job.ok, nblocks);
    memcpy(err, job.err, sizeof(float) * nblocks);
    memcpy(sig, job.sig, sizeof(float) * nblocks);
    if (llr && nblocks <= drx_ldpc_lanes())
        for (int b = 0; b < nblocks; b++)
            memcpy(llr + (size_t)b * w->fec.nldpc, w->llr[b], w->fec.nldpc);
    rc = 0;
out:
    free(job.bb);
    free(job.ok);
    free(job.err);
    free(job.sig);
    if (w) {
        for (int i = 0; i < 64; i++) {
            free(w->llr[i]);
            free(w->cw[i]);
        }
        drx_ldpc_free(w->ldpc);
        drx_bch_free(w->bch);
        free(w->bitmap);
        free(w->cellperm);
        free(w->ci);
        free(w->cq);
        free(w->cc);
        free(w->ck);
        free(w->planes);
        free(w->llrsrc);
        free(w->il);
    }
    free(w);
    free(dd);
    return rc;
}
