/* A plain C99 caller of include/celeritas_amd.h, as an FFI binding would
 * use it: no C++ and no Python. Without a device it exercises the host-only
 * entry points and the error path; with one it runs the SimpleCompton
 * problem (problem JSON from argv[1]) to completion through the stepper
 * and prints the step iterations and the summed active-track counts.
 *
 * Usage: capi_caller g4emlow_dir [problem.json primaries]
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "celeritas_amd.h"

static char* read_file(const char* path)
{
    FILE* f = fopen(path, "rb");
    if (!f)
        return NULL;
    fseek(f, 0, SEEK_END);
    long n = ftell(f);
    fseek(f, 0, SEEK_SET);
    char* buf = (char*)malloc((size_t)n + 1);
    if (fread(buf, 1, (size_t)n, f) != (size_t)n)
    {
        fclose(f);
        free(buf);
        return NULL;
    }
    buf[n] = '\0';
    fclose(f);
    return buf;
}

int main(int argc, char** argv)
{
    if (argc < 2)
    {
        fprintf(stderr, "usage: %s g4emlow_dir [problem.json primaries]\n", argv[0]);
        return 2;
    }
    printf("version %s\n", celer_amd_version());
    printf("sizes %zu %zu %zu\n", sizeof(CelerPrimary), sizeof(CelerStepperResult),
           sizeof(CelerThreadSlots));

    /* Host-only: an element's SB table as JSON (size query, then fill) */
    size_t needed = 0;
    if (celer_read_seltzer_berger(argv[1], 29, NULL, 0, &needed) != 0)
    {
        printf("error %s\n", celer_amd_last_error());
        return 1;
    }
    char* text = (char*)malloc(needed);
    if (celer_read_seltzer_berger(argv[1], 29, text, needed, &needed) != 0)
        return 1;
    printf("sb_json %zu %.12s\n", strlen(text), text);
    free(text);
    /* Errors are return codes with a message */
    if (celer_read_seltzer_berger(argv[1], 101, NULL, 0, &needed) == 0)
        return 1;
    printf("error_message %s\n", celer_amd_last_error());

    if (!celer_amd_device_available())
    {
        CelerParams* none = celer_params_create("{}");
        printf("no_device %d %s\n", none == NULL, celer_amd_last_error());
        return none == NULL ? 0 : 1;
    }
    if (argc < 4)
        return 0;

    char* json = read_file(argv[2]);
    CelerParams* params = json ? celer_params_create(json) : NULL;
    free(json);
    if (!params)
    {
        printf("error %s\n", celer_amd_last_error());
        return 1;
    }
    CelerStepper* stepper = celer_stepper_create(params, 0, 64, 0);
    if (!stepper)
    {
        printf("error %s\n", celer_amd_last_error());
        return 1;
    }
    /* SimpleCompton primaries: 100 MeV photons at (-22, 0, 0) along +x */
    unsigned n = (unsigned)atoi(argv[3]);
    CelerPrimary* prim = (CelerPrimary*)calloc(n, sizeof(CelerPrimary));
    for (unsigned i = 0; i < n; ++i)
    {
        prim[i].particle_id = 0;
        prim[i].event_id = 0;
        prim[i].energy = 100.0;
        prim[i].pos[0] = -22.0;
        prim[i].dir[0] = 1.0;
    }
    CelerStepperResult r;
    unsigned long long active = 0;
    unsigned iters = 0;
    int rc = celer_stepper_step(stepper, prim, n, &r);
    while (rc == 0)
    {
        ++iters;
        active += r.active;
        if (!(r.queued > 0 || r.alive > 0) || iters > 100000)
            break;
        rc = celer_stepper_step(stepper, NULL, 0, &r);
    }
    free(prim);
    if (rc != 0)
    {
        printf("error %s\n", celer_amd_last_error());
        return 1;
    }
    printf("iterations %u active %llu\n", iters, active);
    celer_stepper_destroy(stepper);
    celer_params_destroy(params);
    return 0;
}
