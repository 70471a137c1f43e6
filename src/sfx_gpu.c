/* sfx_gpu.c -- sincFeyn's GPU back-end ABI (reference driver/gpu.h:12-22) implemented over the MI355X
 * library (include/sfx.h): the file a sincFeyn build links in place of driver/gpu.cu and the generated
 * per-graph CUDA launchers, so that its unchanged host driver (driver/sum.c, sumkey.h, eval.c and
 * lattice.c compiled with -DSQ_GPU) runs its batches on the device.  The reference's semantics are kept:
 * SINCQED_GPU=0 turns the device off, SINCQED_GPU=d picks device d-1 (gpu.cu:1-6, 28-39); a graph
 * without a device kernel stays on the CPU (gpu_setup returns 0); two slots, submit returns at once and
 * wait hands out host arrays valid until the slot's next submit; any device error ends the process with
 * exit code 3 (gpu.cu:16). */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "sfx.h"

static int ON = -1;

static void die(const char *what){
    fprintf(stderr, "sfx GPU back end: %s: %s\n", what, sfx_last_error());
    exit(3);
}

static int init(void){
    if(ON >= 0) return ON;
    const char *e = getenv("SINCQED_GPU");
    ON = 0;
    if(e && !strcmp(e, "0")) return 0;
    int dev = e && atoi(e) > 0 ? atoi(e) - 1 : 0;
    if(sfx_init(dev) != SFX_OK && (dev == 0 || sfx_init(0) != SFX_OK)) return 0;     /* no device: the CPU path */
    return ON = 1;
}

const char *gpu_info(void){ return ON == 1 ? sfx_info() : "off"; }

int gpu_setup(const char *name, int D, const double *EX, int K, double H){
    if(!init()) return 0;
    int r = sfx_slot_setup(name, D, EX, K, H);
    if(r < 0) die("gpu_setup");
    return r;
}

void gpu_submit(int slot, long n, const int *k, int stride, const unsigned char *skip){
    if(sfx_slot_submit(slot, n, (const int32_t*)k, stride, skip) != SFX_OK) die("gpu_submit");
}

void gpu_submit_cells(int slot, long nc, const int *cells, int stride, const unsigned char *skip, int st){
    if(sfx_slot_submit_cells(slot, nc, (const int32_t*)cells, stride, skip, st) != SFX_OK) die("gpu_submit_cells");
}

void gpu_wait(int slot, const double **fw, const double **abw){
    if(sfx_slot_wait(slot, fw, abw) != SFX_OK) die("gpu_wait");
}
