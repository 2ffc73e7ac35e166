/*
 * wc_sanitize_main.c -- command-line driver of the CPU restatement and the scene generators, built
 * with AddressSanitizer and UndefinedBehaviorSanitizer by oracle/sanitize.mk.
 *
 * TEST INFRASTRUCTURE ONLY (tests/test_sanitize.py runs it).  Each command reads its inputs from
 * files, runs one piece of host code under the sanitizers and writes the result, so the test can
 * compare it bit for bit with the unsanitized builds; any out-of-bounds access, use after free,
 * leak, signed overflow, misaligned access or out-of-range float-to-int conversion aborts with a
 * report instead.
 *
 *   step      <params.bin> <in.bin> <n> <dt> <steps> <out.bin>   wco_step (sort, density, update)
 *             out: buffer 1, counts, offsets, sorted order
 *   advect    <params.bin> <in.bin> <n> <dt> <reps> <out.bin>    wco_advect
 *   block     nx ny nz ox oy oz spacing jitter vx vy vz seed <out.bin>     wc_scene_block
 *   sphere    count cx cy cz spacing jitter vx vy vz seed <out.bin>       wc_scene_sphere
 *   reference n particle_radius seed <out.bin>                           wc_scene_reference
 * The scene commands print the generator's status code.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../include/watercube.h"

void wco_step(const wc_params* P, wc_particle* buf1, wc_particle* buf2, uint32_t* keys, uint32_t* counts,
              uint32_t* offsets, uint32_t* sorted, float dt, uint32_t nsteps, int nthreads);
void wco_advect(const wc_params* P, wc_particle* ps, uint32_t n, float dt);

static void* read_file(const char* path, size_t bytes) {
    void* buf = calloc(bytes ? bytes : 1, 1);
    FILE* f = fopen(path, "rb");
    if (!buf || !f || (bytes && fread(buf, 1, bytes, f) != bytes)) {
        fprintf(stderr, "cannot read %zu bytes from %s\n", bytes, path);
        exit(2);
    }
    fclose(f);
    return buf;
}

static void write_file(FILE* f, const void* p, size_t bytes) {
    if (bytes && fwrite(p, 1, bytes, f) != bytes) {
        fprintf(stderr, "short write\n");
        exit(2);
    }
}

static int cmd_step(int argc, char** argv, int advect) {
    if (argc != 8) return 2;
    wc_params* P = (wc_params*)read_file(argv[2], sizeof(wc_params));
    const uint32_t n = (uint32_t)strtoul(argv[4], NULL, 10);
    const float dt = strtof(argv[5], NULL);
    const uint32_t steps = (uint32_t)strtoul(argv[6], NULL, 10);
    wc_particle* buf1 = (wc_particle*)read_file(argv[3], (size_t)n * sizeof(wc_particle));
    FILE* out = fopen(argv[7], "wb");
    if (!out) return 2;
    if (advect) {
        for (uint32_t s = 0; s < steps; ++s) wco_advect(P, buf1, n, dt);
        write_file(out, buf1, (size_t)n * sizeof(wc_particle));
    } else {
        const size_t R = (size_t)P->grid_res;
        const size_t cells = R * R * R;
        wc_particle* buf2 = (wc_particle*)malloc((n ? n : 1) * sizeof(wc_particle));
        uint32_t* keys = (uint32_t*)calloc(n ? n : 1, 4);
        uint32_t* sorted = (uint32_t*)calloc(n ? n : 1, 4);
        uint32_t* counts = (uint32_t*)calloc(cells, 4);
        uint32_t* offsets = (uint32_t*)calloc(cells, 4);
        if (n) memcpy(buf2, buf1, (size_t)n * sizeof(wc_particle));
        wco_step(P, buf1, buf2, keys, counts, offsets, sorted, dt, steps, 4);
        write_file(out, buf1, (size_t)n * sizeof(wc_particle));
        write_file(out, counts, cells * 4);
        write_file(out, offsets, cells * 4);
        write_file(out, sorted, (size_t)n * 4);
        free(buf2);
        free(keys);
        free(sorted);
        free(counts);
        free(offsets);
    }
    fclose(out);
    free(buf1);
    free(P);
    return 0;
}

static int cmd_scene(int argc, char** argv) {
    const char* what = argv[1];
    int rc;
    uint32_t n;
    wc_particle* buf;
    const char* path;
    if (!strcmp(what, "block") && argc == 15) {
        const uint32_t nx = (uint32_t)atoi(argv[2]), ny = (uint32_t)atoi(argv[3]), nz = (uint32_t)atoi(argv[4]);
        const float origin[3] = {strtof(argv[5], NULL), strtof(argv[6], NULL), strtof(argv[7], NULL)};
        const float vel[3] = {strtof(argv[10], NULL), strtof(argv[11], NULL), strtof(argv[12], NULL)};
        n = nx * ny * nz;
        buf = (wc_particle*)calloc(n ? n : 1, sizeof(wc_particle));
        rc = wc_scene_block(buf, nx, ny, nz, origin, strtof(argv[8], NULL), strtof(argv[9], NULL), vel,
                            strtoull(argv[13], NULL, 10));
        path = argv[14];
    } else if (!strcmp(what, "sphere") && argc == 13) {
        n = (uint32_t)atoi(argv[2]);
        const float center[3] = {strtof(argv[3], NULL), strtof(argv[4], NULL), strtof(argv[5], NULL)};
        const float vel[3] = {strtof(argv[8], NULL), strtof(argv[9], NULL), strtof(argv[10], NULL)};
        buf = (wc_particle*)calloc(n ? n : 1, sizeof(wc_particle));
        rc = wc_scene_sphere(buf, n, center, strtof(argv[6], NULL), strtof(argv[7], NULL), vel,
                             strtoull(argv[11], NULL, 10));
        path = argv[12];
    } else if (!strcmp(what, "reference") && argc == 6) {
        n = (uint32_t)atoi(argv[2]);
        buf = (wc_particle*)calloc(n ? n : 1, sizeof(wc_particle));
        rc = wc_scene_reference(buf, n, strtof(argv[3], NULL), strtoull(argv[4], NULL, 10));
        path = argv[5];
    } else {
        return 2;
    }
    FILE* out = fopen(path, "wb");
    if (!out) return 2;
    write_file(out, buf, (size_t)n * sizeof(wc_particle));
    fclose(out);
    free(buf);
    printf("%d\n", rc);
    return 0;
}

int main(int argc, char** argv) {
    if (argc < 2) return 2;
    if (!strcmp(argv[1], "step")) return cmd_step(argc, argv, 0);
    if (!strcmp(argv[1], "advect")) return cmd_step(argc, argv, 1);
    return cmd_scene(argc, argv);
}
