/*
 * The physim pipeline loop (physim-core/src/pipeline.rs:143-182) driving the
 * device-resident verlet element through the C-ABI, as the Rust shim would:
 * the GPU gravity element is created through its TransformElementAPI, the
 * verlet element is bound to it, and each iteration runs
 *   synths (state.extend, new_state.extend) -> integrate(state, new_state)
 *   -> transmutes(new_state) -> state = new_state.
 * A synth appends one star at iteration `append_at` (a length change), a
 * transmute moves body 7 by `edit_dx` along x at iteration `edit_at` (an
 * edited state of the same length). The final state is written as raw
 * Entity records.
 *
 *   pipeline_harness N_CUBE STEPS THETA E DT PRECISION EDIT_AT APPEND_AT OUT
 *
 * N_CUBE = 0 runs three free bodies instead of the cube + two stars.
 * Exit status: 0 on success, 2 on a usage error, 1 on an engine error.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "physim_abi.h"
#include "physim_gpu.h"

const TransformElementAPI* astro2_get_api(void);

static int fail(const char* what) {
    fprintf(stderr, "pipeline_harness: %s: %s\n", what, pg_last_error());
    return 1;
}

int main(int argc, char** argv) {
    if (argc != 10) {
        fprintf(stderr, "usage: %s N_CUBE STEPS THETA E DT PRECISION EDIT_AT APPEND_AT OUT\n",
                argv[0]);
        return 2;
    }
    const unsigned long n_cube = strtoul(argv[1], NULL, 10);
    const int steps = atoi(argv[2]);
    const double theta = atof(argv[3]), e = atof(argv[4]), dt = atof(argv[5]);
    const char* precision = argv[6];
    const int edit_at = atoi(argv[7]), append_at = atoi(argv[8]);
    const char* out_path = argv[9];

    size_t n = n_cube ? n_cube + 2 : 3;
    Entity* state = calloc(n + 1, sizeof(Entity));
    Entity* new_state = calloc(n + 1, sizeof(Entity));
    if (!state || !new_state) return 1;
    if (n_cube) {
        const double c0[3] = {0.0, 0.0, 0.0};
        if (pg_gen_cube(n_cube, 1, 1000.0, 1.0, c0, 1.0, 0, state) ||
            pg_gen_star(0.2, 0.2, 0.5, 0, 0, 0, 0.1, 1e5, 0, 0, state + n_cube) ||
            pg_gen_star(-0.2, -0.2, 0.5, 0, 0, 0, 0.1, 1e5, 0, 0, state + n_cube + 1))
            return fail("generators");
    } else {
        pg_gen_star(0.31, 0.12, -0.07, 0.3, -1.0, 2.5, 0.1, 1.7, 0, 0, state);
        pg_gen_star(-0.45, 0.25, 0.5, -0.7, 0.1, 0.0, 0.1, 2.3, 0, 0, state + 1);
        pg_gen_star(0.05, -0.61, 0.2, 0.2, 0.4, -0.3, 0.1, 0.9, 0, 0, state + 2);
    }
    memcpy(new_state, state, n * sizeof(Entity));

    /* the GPU astro2 element, as the host builds it (transform.rs:65-76) */
    char props[256];
    snprintf(props, sizeof props, "{\"theta\": %.17g, \"e\": %.17g, \"precision\": \"%s\"}",
             theta, e, precision);
    const TransformElementAPI* api = astro2_get_api();
    void* element = api->init((const uint8_t*)props, strlen(props));
    if (!element) return fail("astro2 init");

    pg_verlet* v = NULL;
    if (pg_verlet_create(&v) != PG_OK) return fail("verlet create");
    /* unbound: the shim would run the stock verlet with acc_fn */
    if (pg_verlet_integrate(v, state, new_state, n, dt) != PG_ERR_STATE) {
        fprintf(stderr, "pipeline_harness: an unbound verlet must refuse to integrate\n");
        return 1;
    }
    if (pg_verlet_bind(v, element) != PG_OK) return fail("verlet bind");

    for (int it = 1; it <= steps; ++it) {
        if (it == append_at) { /* a synth adds a star to both vectors */
            Entity s;
            pg_gen_star(0.05, -0.3, 0.25, 1.0, 0.0, 0.0, 0.1, 3.0, 0, 0, &s);
            state[n] = s;
            new_state[n] = s;
            ++n;
        }
        if (pg_verlet_integrate(v, state, new_state, n, dt) != PG_OK) return fail("integrate");
        if (it == edit_at) new_state[n > 7 ? 7 : 0].x += 1e-3; /* a transmute */
        memcpy(state, new_state, n * sizeof(Entity));
    }

    FILE* f = fopen(out_path, "wb");
    if (!f || fwrite(state, sizeof(Entity), n, f) != n) return fail("write");
    fclose(f);
    pg_verlet_destroy(v);
    api->destroy(element);
    printf("pipeline_harness: %d iterations of %zu bodies\n", steps, n);
    free(state);
    free(new_state);
    return 0;
}
