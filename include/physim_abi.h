/*
 * physim plugin ABI, as exported by libphysim_gpu.so.
 *
 * These are the symbols the reference host resolves when it scans a plugin
 * directory and builds a pipeline. A GPU `astro`, `astro2` or `simple_astro`
 * element drops in by placing libphysim_gpu.so in PHYSIM_PLUGIN_DIR; no
 * pipeline change is needed (element_db keeps the last plugin discovered for
 * a name: physim-core/src/plugin/discover.rs:256-259,262-281).
 *
 * The types restate the reference's repr(C) definitions
 * (physim-core/src/lib.rs:16-37, physim-core/src/messages.rs:10-45,
 * physim-core/src/plugin/meta.rs:246-257, physim-core/src/plugin/mod.rs:60-69,
 * physim-core/src/plugin/transform.rs:32-50; C mirror c_plugin/physim.h:6-94).
 */
#ifndef PHYSIM_ABI_H
#define PHYSIM_ABI_H

#include <stdbool.h>
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* physim-core/src/plugin/mod.rs:60-69 */
typedef enum PhysimElementKind {
    PHYSIM_INITIALISER = 0,
    PHYSIM_TRANSFORM = 1,
    PHYSIM_RENDER = 2,
    PHYSIM_SYNTH = 3,
    PHYSIM_TRANSMUTE = 4,
    PHYSIM_INTEGRATOR = 5,
} PhysimElementKind;

/* physim-core/src/messages.rs:10-19 */
typedef enum PhysimMessagePriority {
    PHYSIM_PRIO_BACKGROUND = 0,
    PHYSIM_PRIO_LOW = 1,
    PHYSIM_PRIO_NORMAL = 2,
    PHYSIM_PRIO_HIGH = 3,
    PHYSIM_PRIO_REALTIME = 4,
    PHYSIM_PRIO_CRITICAL = 5,
} PhysimMessagePriority;

/* physim-core/src/messages.rs:28-45 (origin 0 = Rust, 1 = C) */
typedef struct CMessage {
    PhysimMessagePriority priority;
    const char* topic;
    const char* message;
    uintptr_t sender_id;
    int origin;
} CMessage;

/* physim-core/src/lib.rs:16-29: 80 bytes, id at 64, fixed at 72. */
typedef struct Entity {
    double x, y, z;
    double vx, vy, vz;
    double radius;
    double mass;
    uintptr_t id;
    bool fixed;
} Entity;

/* physim-core/src/lib.rs:31-37 */
typedef struct Acceleration {
    double x, y, z;
} Acceleration;

/* The layouts the reference's repr(C) structs fix (checked at compile time
 * wherever this header is included). */
#if defined(__cplusplus)
static_assert(sizeof(Entity) == 80 && offsetof(Entity, id) == 64 && offsetof(Entity, fixed) == 72,
              "Entity must match physim-core's repr(C) layout");
static_assert(sizeof(Acceleration) == 24, "Acceleration is three doubles");
#elif defined(__STDC_VERSION__) && __STDC_VERSION__ >= 201112L
_Static_assert(sizeof(Entity) == 80 && offsetof(Entity, id) == 64 && offsetof(Entity, fixed) == 72,
               "Entity must match physim-core's repr(C) layout");
_Static_assert(sizeof(Acceleration) == 24, "Acceleration is three doubles");
#endif

/* physim-core/src/plugin/meta.rs:317 */
typedef char* (*RustStringAllocFn)(const char*);

/* physim-core/src/plugin/transform.rs:32-50 */
typedef struct TransformElementAPI {
    void* (*init)(const uint8_t* json, uintptr_t len);
    void (*transform)(const void* obj, const Entity* state, uintptr_t state_len,
                      Acceleration* acceleration, uintptr_t acceleration_len);
    void (*destroy)(void* obj);
    char* (*get_property_descriptions)(void* obj, RustStringAllocFn alloc);
    void (*recv_message)(void* obj, const CMessage* msg);
    void (*post_configuration_messages)(void* obj);
} TransformElementAPI;

/* physim-core/src/plugin/meta.rs:246-257 */
typedef struct ElementMetaFFI {
    PhysimElementKind kind;
    char* name;
    char* plugin;
    char* version;
    char* license;
    char* author;
    char* blurb;
    char* repo;
} ElementMetaFFI;

/* ---- library-level symbols ------------------------------------------- */

/* Replaces the register_plugin! macro's get_plugin_abi_info
 * (physim-core/src/plugin/mod.rs:124-127; C form c_plugin/plugin.c:39-41).
 * Returns "C": the host accepts it silently (discover.rs:346). */
const char* get_plugin_abi_info(void);

/* Replaces register_plugin (physim-core/src/plugin/mod.rs:129-136; C form
 * c_plugin/plugin.c:43-45): "astro,astro2,simple_astro". */
const char* register_plugin(void);

/* Replaces set_callback_target (physim-core/src/plugin/mod.rs:138-145;
 * C form c_plugin/plugin.c:47-53). Aborts on NULL, as both do. */
void set_callback_target(void* target);

/* ---- per-element symbols (physim-attribute/src/lib.rs:39-202) --------- */

/* astro2: Barnes-Hut octree gravity (astro/src/transformers.rs:114-203). */
ElementMetaFFI astro2_register(RustStringAllocFn alloc);
const TransformElementAPI* astro2_get_api(void);

/* astro: Barnes-Hut quadtree gravity (astro/src/transformers.rs:16-112). */
ElementMetaFFI astro_register(RustStringAllocFn alloc);
const TransformElementAPI* astro_get_api(void);

/* simple_astro: direct O(N^2) gravity (astro/src/transformers.rs:207-271). */
ElementMetaFFI simple_astro_register(RustStringAllocFn alloc);
const TransformElementAPI* simple_astro_get_api(void);

#ifdef __cplusplus
}
#endif

#endif /* PHYSIM_ABI_H */
