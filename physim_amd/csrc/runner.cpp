// A device-resident pipeline runner: Pipeline::run's simulation loop
// (physim-core/src/pipeline.rs:134-192) for a pipeline of one GPU gravity
// transform and an integrator, with the state on the device between
// iterations, and the csvsink renderer (utilities/src/csvsink.rs:50-79) as
// its frame writer. A frame crosses PCIe only when the sink prints it.
//
// Order per iteration, as pipeline.rs:155-182 with no synths or transmutes:
// integrate (acc_fn = the transform), state = new_state, hand the state to
// the renderer. The initial state is handed over before the loop
// (pipeline.rs:128-131). The renderer runs on its own thread behind a queue
// of two frames (the reference's sync_channel(2)), so formatting overlaps
// the device's next steps.
#include <charconv>
#include <cmath>
#include <condition_variable>
#include <cstdio>
#include <cstring>
#include <deque>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "pg_engine.h"
#include "physim_gpu.h"

// Rust's `Display` for f64 (what csvsink's write!("{},{},{},", x, y, z)
// prints): the shortest decimal digits that read back to the same double,
// in positional notation without exponent or trailing ".0"; "NaN", "inf",
// "-inf", and "-0" for negative zero.
extern "C" size_t pg_format_f64(double v, char* out, size_t cap) {
    char buf[400];
    size_t k = 0;
    auto put = [&](char c) {
        if (k < sizeof buf) buf[k++] = c;
    };
    if (v != v) {
        for (const char* s = "NaN"; *s; ++s) put(*s);
    } else if (std::isinf(v)) {
        if (v < 0) put('-');
        for (const char* s = "inf"; *s; ++s) put(*s);
    } else {
        if (std::signbit(v)) {
            put('-');
            v = -v;
        }
        if (v == 0.0) {
            put('0');
        } else {
            // shortest round-trip digits d1.d2...dm and exponent E: v = 0.d1...dm * 10^(E + 1)
            char sci[64];
            auto r = std::to_chars(sci, sci + sizeof sci, v, std::chars_format::scientific);
            char* e = (char*)memchr(sci, 'e', (size_t)(r.ptr - sci));
            char digits[32];
            int m = 0;
            for (char* p = sci; p < e; ++p)
                if (*p != '.') digits[m++] = *p;
            int exp10 = 0;
            std::from_chars(e + 1 + (e[1] == '+' ? 1 : 0), r.ptr, exp10);
            const int point = exp10 + 1;  // digits before the decimal point
            if (point <= 0) {
                put('0');
                put('.');
                for (int i = 0; i < -point; ++i) put('0');
                for (int i = 0; i < m; ++i) put(digits[i]);
            } else if (point >= m) {
                for (int i = 0; i < m; ++i) put(digits[i]);
                for (int i = m; i < point; ++i) put('0');
            } else {
                for (int i = 0; i < point; ++i) put(digits[i]);
                put('.');
                for (int i = point; i < m; ++i) put(digits[i]);
            }
        }
    }
    if (out && cap) {
        const size_t c = k < cap - 1 ? k : cap - 1;
        memcpy(out, buf, c);
        out[c] = 0;
    }
    return k;
}

namespace {

struct CsvSink {
    std::string file;
    uint32_t print_n = 1;
    FILE* f = nullptr;
};

struct Frame {
    uint64_t iteration = 0;  // states handed to the renderer before this one
    std::vector<Entity> state;
};

}  // namespace

struct pg_runner {
    pg_engine* eng = nullptr;
    pg_sim* sim = nullptr;
    uint64_t n = 0;
    uint64_t iterations_done = 0;
    bool sent_initial = false;
    std::vector<CsvSink> sinks;
    // renderer thread and its two-frame queue
    std::thread writer;
    std::mutex mu;
    std::condition_variable cv;
    std::deque<Frame> queue;
    std::vector<std::vector<Entity>> spare;
    bool stop = false;
    bool busy = false;  // the writer holds a frame
    int write_error = 0;
};

// csvsink.rs:73-79: "x,y,z," per entity, then a newline.
static int print_state(FILE* f, const std::vector<Entity>& state) {
    std::string line;
    line.reserve(state.size() * 60 + 1);
    char num[400];
    for (const Entity& e : state) {
        for (double c : {e.x, e.y, e.z}) {
            const size_t k = pg_format_f64(c, num, sizeof num);
            line.append(num, k);
            line.push_back(',');
        }
    }
    line.push_back('\n');
    return fwrite(line.data(), 1, line.size(), f) == line.size() ? 0 : -1;
}

// csvsink.rs:50-71: the first state received is printed, then each later
// state whose count of states received before it is a multiple of print_n.
static bool prints(const CsvSink& s, uint64_t iteration) {
    return iteration == 0 || iteration % s.print_n == 0;
}

static void writer_loop(pg_runner* r) {
    for (;;) {
        Frame frame;
        {
            std::unique_lock<std::mutex> lk(r->mu);
            r->cv.wait(lk, [&] { return r->stop || !r->queue.empty(); });
            if (r->queue.empty()) return;
            frame = std::move(r->queue.front());
            r->queue.pop_front();
            r->busy = true;
        }
        r->cv.notify_all();
        int err = 0;
        for (CsvSink& s : r->sinks)
            if (prints(s, frame.iteration) && print_state(s.f, frame.state) != 0) err = 1;
        {
            std::lock_guard<std::mutex> lk(r->mu);
            r->write_error |= err;
            r->spare.push_back(std::move(frame.state));
            r->busy = false;
        }
        r->cv.notify_all();
    }
}

// Whether any sink prints the state with this iteration count.
static bool wanted(const pg_runner* r, uint64_t iteration) {
    for (const CsvSink& s : r->sinks)
        if (prints(s, iteration)) return true;
    return false;
}

// Downloads the current state and queues it for the renderer (blocking while
// two frames wait, as a full sync_channel(2) blocks the simulation).
static int emit(pg_runner* r, uint64_t iteration) {
    Frame frame;
    frame.iteration = iteration;
    {
        std::unique_lock<std::mutex> lk(r->mu);
        r->cv.wait(lk, [&] { return r->queue.size() < 2; });
        if (!r->spare.empty()) {
            frame.state.swap(r->spare.back());
            r->spare.pop_back();
        }
    }
    frame.state.resize(r->n);
    int rc = pg_sim_download(r->sim, frame.state.data(), r->n);
    if (rc != PG_OK) return rc;
    {
        std::lock_guard<std::mutex> lk(r->mu);
        r->queue.push_back(std::move(frame));
    }
    r->cv.notify_all();
    return PG_OK;
}

extern "C" int pg_runner_create(pg_engine* eng, const pg_params* params, int32_t integrator,
                                const Entity* initial, uint64_t n, pg_runner** out) {
    if (!eng || !params || !out || (n && !initial))
        return pg_set_error(PG_ERR_ARG, "NULL argument");
    pg_runner* r = new pg_runner();
    r->eng = eng;
    r->n = n;
    int rc = pg_sim_create(eng, n, params, &r->sim);
    if (rc == PG_OK && integrator != PG_INTEG_VERLET) rc = pg_sim_set_integrator(r->sim, integrator);
    if (rc == PG_OK) rc = pg_sim_upload(r->sim, initial, n);
    if (rc != PG_OK) {
        if (r->sim) pg_sim_destroy(r->sim);
        delete r;
        return rc;
    }
    *out = r;
    return PG_OK;
}

extern "C" int pg_runner_add_csvsink(pg_runner* r, const char* file, uint32_t print_n) {
    if (!r) return pg_set_error(PG_ERR_ARG, "runner is NULL");
    if (r->sent_initial) return pg_set_error(PG_ERR_STATE, "sinks are added before the run");
    CsvSink s;
    s.file = file && *file ? file : "csvsink.csv";  // csvsink.rs:35-38 defaults
    s.print_n = print_n ? print_n : 1;
    // csvsink.rs:42-48: created or truncated when the renderer starts
    s.f = fopen(s.file.c_str(), "w");
    if (!s.f) return pg_set_error(PG_ERR_ARG, "cannot open %s", s.file.c_str());
    r->sinks.push_back(s);
    return PG_OK;
}

extern "C" int pg_runner_run(pg_runner* r, uint32_t iterations) {
    if (!r) return pg_set_error(PG_ERR_ARG, "runner is NULL");
    if (!r->sinks.empty() && !r->writer.joinable()) r->writer = std::thread(writer_loop, r);
    int rc = PG_OK;
    if (!r->sent_initial) {
        r->sent_initial = true;
        if (!r->sinks.empty()) rc = emit(r, 0);  // pipeline.rs:128-131
    }
    // the device steps in runs up to the next state a sink prints; the states
    // in between are counted by the sinks but never leave the device
    uint32_t done = 0;
    while (rc == PG_OK && done < iterations) {
        uint32_t k = iterations - done;
        if (!r->sinks.empty()) {
            k = 1;
            while (done + k < iterations && !wanted(r, r->iterations_done + k)) ++k;
        }
        rc = pg_sim_step(r->sim, k);
        if (rc != PG_OK) break;
        done += k;
        r->iterations_done += k;
        if (wanted(r, r->iterations_done) && !r->sinks.empty()) rc = emit(r, r->iterations_done);
    }
    return rc;
}

// Waits until the renderer has written every frame handed to it.
extern "C" int pg_runner_flush(pg_runner* r) {
    if (!r) return pg_set_error(PG_ERR_ARG, "runner is NULL");
    {
        std::unique_lock<std::mutex> lk(r->mu);
        r->cv.wait(lk, [&] { return r->queue.empty() && !r->busy; });
    }
    for (CsvSink& s : r->sinks)
        if (s.f) fflush(s.f);
    return r->write_error ? pg_set_error(PG_ERR_ARG, "a csvsink write failed") : PG_OK;
}

extern "C" int pg_runner_state(pg_runner* r, Entity* out, uint64_t n) {
    if (!r) return pg_set_error(PG_ERR_ARG, "runner is NULL");
    return pg_sim_download(r->sim, out, n);
}

extern "C" void pg_runner_destroy(pg_runner* r) {
    if (!r) return;
    {
        std::lock_guard<std::mutex> lk(r->mu);
        r->stop = true;
    }
    r->cv.notify_all();
    if (r->writer.joinable()) r->writer.join();
    for (CsvSink& s : r->sinks)
        if (s.f) fclose(s.f);
    if (r->sim) pg_sim_destroy(r->sim);
    delete r;
}
