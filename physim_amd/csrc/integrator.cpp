// The `verlet` integrator element with its state on the device: the C side
// of a Rust shim that keeps the reference element's name and interface
// (IntegratorElement::integrate, physim-core/src/plugin/integrator.rs:5-13;
// Verlet, integrators/src/verlet.rs:85-108). The shim binds it to the GPU
// gravity element that announced itself on the message bus; each integrate
// call then steps on the device and returns new_state.
//
// The device keeps its state between calls. The pipeline hands back, as the
// next call's input, the new_state the last call returned -- unless a synth
// or transmute changed it (pipeline.rs:155-171). Every field of every entity
// is compared with what the last call returned: if they are all equal the
// device state is current; otherwise the input is uploaded, with the last
// call's input as the Stormer history when the length is unchanged (the
// reference restarts with its Taylor step only on a length change,
// verlet.rs:102-107).
#include <cstring>
#include <vector>

#include "pg_engine.h"
#include "physim_gpu.h"

struct pg_verlet {
    const void* element = nullptr;  // the bound GPU gravity element
    pg_sim* sim = nullptr;
    uint64_t n = 0;
    std::vector<Entity> last_in, last_out;
    bool started = false;  // a step was taken (the reference's previous_state is set)
};

extern "C" int pg_verlet_create(pg_verlet** out) {
    if (!out) return pg_set_error(PG_ERR_ARG, "out is NULL");
    *out = new pg_verlet();
    return PG_OK;
}

extern "C" void pg_verlet_destroy(pg_verlet* v) {
    if (!v) return;
    if (v->sim) pg_sim_destroy(v->sim);
    delete v;
}

extern "C" int pg_verlet_bind(pg_verlet* v, const void* element) {
    if (!v) return pg_set_error(PG_ERR_ARG, "verlet is NULL");
    pg_params p;
    if (!pg_plugin_element_params(element, &p))
        return pg_set_error(PG_ERR_ARG, "not a GPU gravity element of this library");
    v->element = element;
    return PG_OK;
}

// Bitwise equality of every field (padding excluded; NaNs compare by bits).
static bool same_entity(const Entity& a, const Entity& b) {
    return memcmp(&a, &b, offsetof(Entity, fixed)) == 0 && a.fixed == b.fixed;
}

extern "C" int pg_verlet_integrate(pg_verlet* v, const Entity* entities, Entity* new_state,
                                   uint64_t n, double dt) {
    if (!v) return pg_set_error(PG_ERR_ARG, "verlet is NULL");
    if (n && (!entities || !new_state)) return pg_set_error(PG_ERR_ARG, "state is NULL");
    pg_params p;
    if (!v->element || !pg_plugin_element_params(v->element, &p))
        return pg_set_error(PG_ERR_STATE, "no GPU gravity element is bound");
    p.dt = dt;
    if (v->sim && v->n != n) {
        pg_sim_destroy(v->sim);
        v->sim = nullptr;
    }
    if (!v->sim) {
        int rc = pg_sim_create(pg_plugin_engine(), n, &p, &v->sim);
        if (rc != PG_OK) return rc;
        v->n = n;
    }
    int rc = pg_sim_set_params(v->sim, &p);
    if (rc != PG_OK) return rc;
    if (!v->started || v->last_in.size() != n) {
        // first call, or a length change: the Taylor first step
        // (verlet.rs:102-103, previous_state.len() != entities.len())
        rc = pg_sim_upload(v->sim, entities, n);
    } else {
        bool current = true;
        for (uint64_t i = 0; i < n && current; ++i) current = same_entity(entities[i], v->last_out[i]);
        // an edited state keeps its history: x_prev = the last call's input
        if (!current) rc = pg_sim_upload_state(v->sim, entities, v->last_in.data(), n);
    }
    if (rc != PG_OK) return rc;
    if ((rc = pg_sim_step(v->sim, 1)) != PG_OK) return rc;
    if ((rc = pg_sim_download(v->sim, new_state, n)) != PG_OK) return rc;
    v->last_in.assign(entities, entities + n);
    v->last_out.assign(new_state, new_state + n);
    v->started = true;
    return PG_OK;
}
