/*
 * replay_trajectory.c — a plain C99 caller of the diomedes_hip C ABI.
 *
 * Reads a scene and trajectory file (the DIOTRAJ format of
 * diomedes_amd/trajectory.py: a 64-byte header, the scene as the ABI's own
 * records, then frames), uploads the scene with dio_world_upload or
 * dio_batch_upload, steps it with dio_step (contact lag 1) to each frame's
 * step count and compares the downloaded poses, velocities and broad-phase
 * pairs with the frame bit for bit.  Then it checks the status contract:
 * DIO_INVALID with a message for zero iterations, DIO_CAPACITY with the
 * required size for a pair buffer that is too small, DIO_NOT_READY before an
 * upload.
 *
 *   replay_trajectory [--parse-only] [--mode M] [--device D] FILE
 *
 * --parse-only reads and checks the file and prints its summary without any
 * GPU call.  Exit status: 0 all frames identical, 1 a frame differs, 2 an
 * ABI call failed unexpectedly, 3 a bad file or usage.
 */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "diomedes_hip.h"

typedef struct {
    uint8_t* data;
    size_t size;
    size_t off;
} reader;

static int take(reader* r, void* dst, size_t bytes) {
    if (r->off + bytes > r->size) return 0;
    if (dst) memcpy(dst, r->data + r->off, bytes);
    r->off += bytes;
    return 1;
}

static uint32_t u32_at(const uint8_t* p) {
    return (uint32_t)p[0] | (uint32_t)p[1] << 8 | (uint32_t)p[2] << 16 | (uint32_t)p[3] << 24;
}

typedef struct {
    uint32_t n, c, h, s, flags;
    dio_avbd_options options;
    dio_gpu_body* bodies;
    dio_avbd_body_state* state;
    dio_avbd_constraint* constraints;
    float* hull;
    uint32_t* offsets;
    dio_container* containers;
    size_t frames_off; /* byte offset of the first frame */
} scene_file;

static int fail_file(const char* what) {
    fprintf(stderr, "bad trajectory file: %s\n", what);
    return 0;
}

static void* alloc_n(size_t count, size_t size) {
    return calloc(count ? count : 1, size);
}

static int read_scene(reader* r, scene_file* sf) {
    uint8_t head[64];
    uint32_t sizes[5];
    int k;
    if (!take(r, head, sizeof head)) return fail_file("too short");
    if (memcmp(head, "DIOTRAJ", 8) != 0) return fail_file("bad magic");
    if (u32_at(head + 8) != 1u) return fail_file("unsupported version");
    sf->n = u32_at(head + 12);
    sf->c = u32_at(head + 16);
    sf->h = u32_at(head + 20);
    sf->s = u32_at(head + 24);
    sf->flags = u32_at(head + 28);
    for (k = 0; k < 5; ++k) sizes[k] = u32_at(head + 32 + 4 * k);
    if (sizes[0] != sizeof(dio_gpu_body) || sizes[1] != sizeof(dio_avbd_body_state) ||
        sizes[2] != sizeof(dio_avbd_constraint) || sizes[3] != sizeof(dio_avbd_options) ||
        sizes[4] != sizeof(dio_container))
        return fail_file("record sizes differ from the header's");
    if (sf->s < 1) return fail_file("no scene");
    sf->bodies = alloc_n(sf->n, sizeof *sf->bodies);
    sf->state = alloc_n(sf->n, sizeof *sf->state);
    sf->constraints = alloc_n(sf->c, sizeof *sf->constraints);
    sf->hull = alloc_n((size_t)sf->h * 3, sizeof *sf->hull);
    sf->offsets = alloc_n((size_t)sf->s + 1, sizeof *sf->offsets);
    sf->containers = alloc_n(sf->s, sizeof *sf->containers);
    if (!sf->bodies || !sf->state || !sf->constraints || !sf->hull || !sf->offsets || !sf->containers)
        return fail_file("out of memory");
    if (!take(r, &sf->options, sizeof sf->options) || !take(r, sf->bodies, (size_t)sf->n * sizeof *sf->bodies) ||
        !take(r, sf->state, (size_t)sf->n * sizeof *sf->state) ||
        !take(r, sf->constraints, (size_t)sf->c * sizeof *sf->constraints) ||
        !take(r, sf->hull, (size_t)sf->h * 3 * sizeof *sf->hull) ||
        !take(r, sf->offsets, ((size_t)sf->s + 1) * sizeof *sf->offsets) ||
        !take(r, sf->containers, (size_t)sf->s * sizeof *sf->containers))
        return fail_file("truncated scene");
    if (sf->offsets[0] != 0 || sf->offsets[sf->s] != sf->n) return fail_file("bad scene offsets");
    for (k = 0; k < (int)sf->s; ++k)
        if (sf->offsets[k] > sf->offsets[k + 1]) return fail_file("bad scene offsets");
    sf->frames_off = r->off;
    return 1;
}

typedef struct {
    uint64_t step;
    uint32_t pairs;
    const float* pos; /* n x 3 */
    const float* rot; /* n x 4 */
    const float* vel; /* n x 3 or null */
    const float* ang;
    const uint32_t* pair_list; /* pairs x 2 or null */
} frame_view;

/* Frame payloads are 4-byte aligned in the file (every field is a multiple
   of 4 bytes and the buffer comes from malloc), so they are read in place. */
static int next_frame(reader* r, const scene_file* sf, frame_view* f) {
    uint8_t head[16];
    const size_t n = sf->n;
    if (!take(r, head, sizeof head)) return fail_file("truncated frame header");
    if (memcmp(head, "FRAM", 4) != 0) return fail_file("bad frame tag");
    f->pairs = u32_at(head + 4);
    f->step = (uint64_t)u32_at(head + 8) | (uint64_t)u32_at(head + 12) << 32;
    f->pos = (const float*)(r->data + r->off);
    if (!take(r, NULL, n * 12)) return fail_file("truncated frame");
    f->rot = (const float*)(r->data + r->off);
    if (!take(r, NULL, n * 16)) return fail_file("truncated frame");
    f->vel = f->ang = NULL;
    if (sf->flags & 1u) {
        f->vel = (const float*)(r->data + r->off);
        if (!take(r, NULL, n * 12)) return fail_file("truncated frame");
        f->ang = (const float*)(r->data + r->off);
        if (!take(r, NULL, n * 12)) return fail_file("truncated frame");
    }
    f->pair_list = NULL;
    if (sf->flags & 2u) {
        f->pair_list = (const uint32_t*)(r->data + r->off);
        if (!take(r, NULL, (size_t)f->pairs * 8)) return fail_file("truncated frame");
    }
    return 1;
}

/* First differing element of two float arrays, compared as bit patterns
   (so -0 differs from +0 and NaNs compare by payload); -1 if identical. */
static long first_diff(const float* a, const float* b, size_t count) {
    size_t i;
    if (memcmp(a, b, count * sizeof(float)) == 0) return -1;
    for (i = 0; i < count; ++i)
        if (memcmp(a + i, b + i, sizeof(float)) != 0) return (long)i;
    return -1;
}

static int check_field(const char* what, uint64_t step, const float* got, const float* want, size_t n,
                       size_t width) {
    long d = first_diff(got, want, n * width);
    if (d < 0) return 1;
    fprintf(stderr, "step %llu: %s differs at body %ld component %ld: got %.9g, want %.9g\n",
            (unsigned long long)step, what, d / (long)width, d % (long)width, (double)got[d], (double)want[d]);
    return 0;
}

#define CALL(ctx, expr)                                                                          \
    do {                                                                                         \
        dio_status st_ = (expr);                                                                 \
        if (st_ != DIO_OK) {                                                                     \
            fprintf(stderr, "%s failed: status %d: %s\n", #expr, (int)st_, dio_last_error(ctx)); \
            return 2;                                                                            \
        }                                                                                        \
    } while (0)

/* The status contract, on the stepped world and on a fresh context. */
static int check_statuses(dio_ctx* ctx, const scene_file* sf, uint32_t last_pairs, int device) {
    dio_avbd_options bad = sf->options;
    dio_step_opts so = {1u, 0u, 0u};
    dio_gpu_pair one;
    uint32_t np = 0;
    dio_limits lim;
    dio_ctx* fresh = NULL;
    dio_status st;
    float pos[3];

    bad.iterations = 0;
    st = dio_step(ctx, &bad, &so, 1);
    if (st != DIO_INVALID || dio_last_error(ctx) == NULL || dio_last_error(ctx)[0] == '\0') {
        fprintf(stderr, "zero iterations: expected DIO_INVALID with a message, got %d\n", (int)st);
        return 2;
    }
    printf("zero iterations -> DIO_INVALID (%s)\n", dio_last_error(ctx));
    if (last_pairs > 1) {
        st = dio_world_pairs(ctx, &one, 1, &np);
        if (st != DIO_CAPACITY || np != last_pairs || dio_required_capacity(ctx) != last_pairs) {
            fprintf(stderr, "small pair buffer: expected DIO_CAPACITY for %u pairs, got %d (n %u, required %llu)\n",
                    last_pairs, (int)st, np, (unsigned long long)dio_required_capacity(ctx));
            return 2;
        }
        printf("pair buffer of 1 -> DIO_CAPACITY, %u required\n", np);
    }
    memset(&lim, 0, sizeof lim);
    lim.max_bodies = 4;
    CALL(NULL, dio_create(device, &lim, &fresh));
    st = dio_world_download(fresh, pos, NULL, NULL, NULL, NULL);
    dio_destroy(fresh);
    if (st != DIO_NOT_READY) {
        fprintf(stderr, "download before upload: expected DIO_NOT_READY, got %d\n", (int)st);
        return 2;
    }
    printf("download before upload -> DIO_NOT_READY\n");
    return 0;
}

static int replay(reader* r, const scene_file* sf, uint32_t mode, int device) {
    const size_t n = sf->n;
    dio_limits lim;
    dio_ctx* ctx = NULL;
    dio_step_opts so;
    float *pos, *rot, *vel, *ang;
    dio_gpu_pair* pairs = NULL;
    uint32_t pairs_cap = 0, last_pairs = 0, frames = 0;
    uint64_t done = 0;
    int rc = 0;

    memset(&lim, 0, sizeof lim);
    lim.max_bodies = sf->n;
    lim.max_constraints = sf->c;
    lim.max_hull_vertices = sf->h;
    CALL(NULL, dio_create(device, &lim, &ctx));
    if (sf->h) CALL(ctx, dio_world_upload_hulls(ctx, sf->hull, sf->h));
    if (sf->s == 1)
        CALL(ctx, dio_world_upload(ctx, sf->bodies, sf->state, sf->n, sf->constraints, sf->c,
                                   sf->containers[0].enabled ? &sf->containers[0] : NULL));
    else
        CALL(ctx, dio_batch_upload(ctx, sf->bodies, sf->state, sf->n, sf->constraints, sf->c, sf->offsets, sf->s,
                                   sf->containers));
    pos = alloc_n(n * 3, sizeof(float));
    rot = alloc_n(n * 4, sizeof(float));
    vel = alloc_n(n * 3, sizeof(float));
    ang = alloc_n(n * 3, sizeof(float));
    if (!pos || !rot || !vel || !ang) {
        fprintf(stderr, "out of memory\n");
        return 2;
    }
    so.contact_lag = 1;
    so.mode = mode;
    so.profile = 0;
    while (r->off < r->size && rc == 0) {
        frame_view f;
        if (!next_frame(r, sf, &f)) return 3;
        if (f.step < done) {
            fprintf(stderr, "frame steps go backwards (%llu after %llu)\n", (unsigned long long)f.step,
                    (unsigned long long)done);
            return 3;
        }
        if (f.step > done) CALL(ctx, dio_step(ctx, &sf->options, &so, (uint32_t)(f.step - done)));
        done = f.step;
        CALL(ctx, dio_world_download(ctx, pos, rot, vel, ang, NULL));
        if (!check_field("position", f.step, pos, f.pos, n, 3) || !check_field("orientation", f.step, rot, f.rot, n, 4))
            rc = 1;
        if (rc == 0 && f.vel &&
            (!check_field("velocity", f.step, vel, f.vel, n, 3) ||
             !check_field("angular velocity", f.step, ang, f.ang, n, 3)))
            rc = 1;
        if (rc == 0 && f.pair_list) {
            uint32_t np = 0;
            if (f.pairs > pairs_cap) {
                free(pairs);
                pairs_cap = f.pairs;
                pairs = alloc_n(pairs_cap, sizeof *pairs);
                if (!pairs) {
                    fprintf(stderr, "out of memory\n");
                    return 2;
                }
            }
            CALL(ctx, dio_world_pairs(ctx, pairs, pairs_cap, &np));
            if (np != f.pairs || (np && memcmp(pairs, f.pair_list, (size_t)np * 8) != 0)) {
                fprintf(stderr, "step %llu: pair list differs (%u pairs, want %u)\n", (unsigned long long)f.step, np,
                        f.pairs);
                rc = 1;
            }
            last_pairs = np;
        }
        ++frames;
    }
    if (rc == 0) {
        printf("replayed %u frames to step %llu: poses%s%s bit-identical\n", frames, (unsigned long long)done,
               sf->flags & 1u ? ", velocities" : "", sf->flags & 2u ? ", pairs" : "");
        rc = check_statuses(ctx, sf, last_pairs, device);
    }
    free(pos);
    free(rot);
    free(vel);
    free(ang);
    free(pairs);
    dio_destroy(ctx);
    return rc;
}

static int usage(void) {
    fprintf(stderr, "usage: replay_trajectory [--parse-only] [--mode M] [--device D] FILE\n");
    return 3;
}

int main(int argc, char** argv) {
    const char* path = NULL;
    int parse_only = 0, device = 0, i, rc;
    uint32_t mode = 0;
    FILE* fp;
    long size;
    reader r;
    scene_file sf;

    for (i = 1; i < argc; ++i) {
        if (strcmp(argv[i], "--parse-only") == 0) parse_only = 1;
        else if (strcmp(argv[i], "--mode") == 0 && i + 1 < argc) mode = (uint32_t)strtoul(argv[++i], NULL, 10);
        else if (strcmp(argv[i], "--device") == 0 && i + 1 < argc) device = atoi(argv[++i]);
        else if (argv[i][0] == '-' || path) return usage();
        else path = argv[i];
    }
    if (!path) return usage();
    fp = fopen(path, "rb");
    if (!fp) {
        perror(path);
        return 3;
    }
    fseek(fp, 0, SEEK_END);
    size = ftell(fp);
    fseek(fp, 0, SEEK_SET);
    r.data = malloc(size > 0 ? (size_t)size : 1);
    r.size = size > 0 ? (size_t)size : 0;
    r.off = 0;
    if (!r.data || fread(r.data, 1, r.size, fp) != r.size) {
        fclose(fp);
        fprintf(stderr, "cannot read %s\n", path);
        return 3;
    }
    fclose(fp);
    memset(&sf, 0, sizeof sf);
    if (!read_scene(&r, &sf)) return 3;
    printf("ABI %u; scene: %u bodies, %u constraints, %u hull vertices, %u scenes, flags %u, dt %g, %u iterations\n",
           dio_abi_version(), sf.n, sf.c, sf.h, sf.s, sf.flags, (double)sf.options.dt, sf.options.iterations);
    if (parse_only) {
        uint32_t frames = 0;
        uint64_t last = 0;
        while (r.off < r.size) {
            frame_view f;
            if (!next_frame(&r, &sf, &f)) return 3;
            last = f.step;
            ++frames;
        }
        printf("%u frames, last step %llu\n", frames, (unsigned long long)last);
        rc = 0;
    } else {
        rc = replay(&r, &sf, mode, device);
    }
    free(r.data);
    free(sf.bodies);
    free(sf.state);
    free(sf.constraints);
    free(sf.hull);
    free(sf.offsets);
    free(sf.containers);
    return rc;
}
