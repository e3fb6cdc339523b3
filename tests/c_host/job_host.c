/* A native host of libmtx's context API, written against include/mtx.h only:
 * it follows the job flow of the reference's native/examples/job-host.c:28-103
 * (asynchronous jobs polled from the calling thread, draining cancellation,
 * resume, two contexts agreeing image for image) and writes the three images
 * for the caller to compare with the oracle.
 *
 *   job_host <descriptor> <output dir> <samples>
 *
 * Exit 0 and "PASS <n> checks" on success; 1 names the failed check. */
#define _POSIX_C_SOURCE 199309L
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "mtx.h"

_Static_assert(sizeof(mtx_job_snapshot) == 24, "job snapshot layout");
_Static_assert(sizeof(mtx_context_snapshot) == 24, "context snapshot layout");

static char error[1024];
static unsigned checks;
#define CHECK(test)                                                               \
  do {                                                                            \
    checks++;                                                                     \
    if (!(test)) {                                                                \
      fprintf(stderr, "line %d: %s (last error: %s)\n", __LINE__, #test, error); \
      return 1;                                                                   \
    }                                                                             \
  } while (0)
#define OK(call) CHECK((call) == MTX_OK)
#define STATUS(code, call) CHECK((call) == (code))

static unsigned char *slurp(const char *path, size_t *size) {
  FILE *f = fopen(path, "rb");
  if (!f) return NULL;
  if (fseek(f, 0, SEEK_END) != 0) { fclose(f); return NULL; }
  long n = ftell(f);
  if (n <= 0 || fseek(f, 0, SEEK_SET) != 0) { fclose(f); return NULL; }
  unsigned char *p = malloc((size_t)n);
  if (p && fread(p, 1, (size_t)n, f) != (size_t)n) { free(p); p = NULL; }
  fclose(f);
  *size = (size_t)n;
  return p;
}

static void nap(void) {
  struct timespec t = {0, 200000};
  nanosleep(&t, NULL);
}

static int image(mtx_handle c, uint32_t kind, unsigned char *out, size_t size) {
  size_t written = 0;
  return mtx_context_read_pixels(c, kind, out, size, &written, error, sizeof error) != MTX_OK || written != size;
}

/* Polls until the job leaves the running/cancelling states. */
static int finish(mtx_handle job, uint32_t budget, mtx_job_snapshot *info) {
  for (unsigned i = 0; i < 200000; i++) {
    if (mtx_job_poll(job, budget, error, sizeof error) != MTX_OK) return 1;
    if (mtx_job_info(job, info, error, sizeof error) != MTX_OK) return 1;
    if (info->state == MTX_JOB_COMPLETE || info->state == MTX_JOB_CANCELLED) return 0;
    nap();
  }
  return 1;
}

int main(int argc, char **argv) {
  if (argc != 4) return 2;
  size_t length = 0;
  unsigned char *desc = slurp(argv[1], &length);
  if (!desc || length < 128) return 2;
  uint32_t width, height;
  memcpy(&width, desc + 8, 4);
  memcpy(&height, desc + 12, 4);
  const size_t size = (size_t)width * height * 16;
  const unsigned samples = (unsigned)strtoul(argv[3], NULL, 10);
  if (samples < 8 || samples > 4096) return 2;
  unsigned char *pa = malloc(size), *pb = malloc(size);
  if (!pa || !pb) return 2;

  mtx_render_settings settings;
  STATUS(MTX_INVALID_ARGUMENT, mtx_render_settings_default(NULL));
  OK(mtx_render_settings_default(&settings));
  CHECK(settings.struct_size == 64 && settings.seed == 12345 && settings.max_samples == 64 && settings.bounces == 6 &&
        settings.direct_light_candidates == 2 && settings.bevel_samples == 8 &&
        settings.transmissive_output_alpha_min == 1.0);
  settings.seed = 0;
  settings.bounces = 8;
  settings.direct_light_candidates = 1;
  settings.max_samples = samples;

  mtx_handle a = 0, b = 0, old = 0, ja = 0, jb = 0;
  mtx_job_snapshot info = {.struct_size = sizeof info};
  mtx_context_snapshot cinfo = {.struct_size = sizeof cinfo};
  OK(mtx_context_create(&a, error, sizeof error));
  OK(mtx_context_create(&b, error, sizeof error));
  size_t written = 1;
  STATUS(MTX_INVALID_STATE, mtx_context_read_pixels(a, 0, NULL, 0, &written, error, sizeof error));
  STATUS(MTX_INVALID_STATE, mtx_context_render(a, 1, error, sizeof error));
  OK(mtx_context_commit_with_settings(a, desc, length, &settings, error, sizeof error));
  OK(mtx_context_commit_with_settings(b, desc, length, &settings, error, sizeof error));
  OK(mtx_context_info(a, &cinfo, error, sizeof error));
  CHECK(cinfo.state == MTX_CONTEXT_COMMITTED && cinfo.width == width && cinfo.height == height &&
        cinfo.samples == 0 && cinfo.last_commit == MTX_COMMIT_FULL);
  STATUS(MTX_INVALID_STATE, mtx_context_read_pixels(a, 0, NULL, 0, &written, error, sizeof error));

  /* Argument and handle validation. */
  STATUS(MTX_INVALID_HANDLE, mtx_job_info(0, &info, error, sizeof error));
  STATUS(MTX_INVALID_ARGUMENT, mtx_context_submit(a, 1, NULL, error, sizeof error));
  STATUS(MTX_INVALID_ARGUMENT, mtx_context_submit(a, 0, &old, error, sizeof error));
  STATUS(MTX_INVALID_ARGUMENT, mtx_context_submit(a, 16385, &old, error, sizeof error));

  /* A queued job blocks the context and cancels at once. */
  OK(mtx_context_submit(a, samples, &old, error, sizeof error));
  OK(mtx_job_info(old, &info, error, sizeof error));
  CHECK(info.state == MTX_JOB_QUEUED && info.submitted_samples == 0 && info.completed_samples == 0 &&
        info.base_samples == 0 && info.requested_samples == samples);
  OK(mtx_context_info(a, &cinfo, error, sizeof error));
  CHECK(cinfo.state == MTX_CONTEXT_JOB_ACTIVE);
  STATUS(MTX_INVALID_STATE, mtx_context_render(a, 1, error, sizeof error));
  STATUS(MTX_INVALID_STATE, mtx_context_reset(a, error, sizeof error));
  STATUS(MTX_INVALID_STATE, mtx_context_destroy(a, error, sizeof error));
  STATUS(MTX_INVALID_STATE, mtx_context_read_pixels(a, 0, NULL, 0, &written, error, sizeof error));
  STATUS(MTX_INVALID_STATE, mtx_context_submit(a, 1, &ja, error, sizeof error));
  STATUS(MTX_INVALID_ARGUMENT, mtx_job_poll(old, 65, error, sizeof error));
  OK(mtx_job_poll(old, 0, error, sizeof error));
  OK(mtx_job_info(old, &info, error, sizeof error));
  CHECK(info.submitted_samples == 0);
  OK(mtx_job_cancel(old, error, sizeof error));
  OK(mtx_job_cancel(old, error, sizeof error));
  OK(mtx_job_info(old, &info, error, sizeof error));
  CHECK(info.state == MTX_JOB_CANCELLED && info.completed_samples == 0);

  /* A running job cancelled with a batch in flight drains it. */
  OK(mtx_context_submit(a, samples, &ja, error, sizeof error));
  CHECK(ja != old);
  STATUS(MTX_INVALID_HANDLE, mtx_job_info(old, &info, error, sizeof error));
  OK(mtx_job_poll(ja, 3, error, sizeof error));
  OK(mtx_job_info(ja, &info, error, sizeof error));
  CHECK(info.state == MTX_JOB_RUNNING && info.submitted_samples == 3 && info.completed_samples <= 3);
  OK(mtx_job_cancel(ja, error, sizeof error));
  OK(mtx_job_info(ja, &info, error, sizeof error));
  CHECK(info.state == MTX_JOB_CANCELLING || info.state == MTX_JOB_CANCELLED);
  CHECK(finish(ja, 0, &info) == 0);
  CHECK(info.state == MTX_JOB_CANCELLED && info.completed_samples == 3 && info.submitted_samples == 3);
  OK(mtx_context_info(a, &cinfo, error, sizeof error));
  CHECK(cinfo.state == MTX_CONTEXT_COMMITTED && cinfo.samples == 3);
  OK(mtx_context_render(b, 3, error, sizeof error));
  for (uint32_t k = 0; k < 3; k++) {
    CHECK(image(a, k, pa, size) == 0 && image(b, k, pb, size) == 0);
    CHECK(memcmp(pa, pb, size) == 0);
  }

  /* Resumed jobs on both contexts, polled with different budgets. */
  old = ja;
  OK(mtx_context_submit(a, samples - 3, &ja, error, sizeof error));
  STATUS(MTX_INVALID_HANDLE, mtx_job_info(old, &info, error, sizeof error));
  OK(mtx_context_submit(b, samples - 3, &jb, error, sizeof error));
  mtx_job_snapshot other = {.struct_size = sizeof other};
  unsigned it = 0;
  for (; it < 200000; it++) {
    OK(mtx_job_poll(ja, 8, error, sizeof error));
    OK(mtx_job_poll(jb, 5, error, sizeof error));
    OK(mtx_job_info(ja, &info, error, sizeof error));
    OK(mtx_job_info(jb, &other, error, sizeof error));
    CHECK(info.base_samples == 3 && info.completed_samples <= info.submitted_samples &&
          info.submitted_samples <= samples - 3);
    CHECK(other.base_samples == 3 && other.completed_samples <= other.submitted_samples);
    if (info.state == MTX_JOB_COMPLETE && other.state == MTX_JOB_COMPLETE) break;
    CHECK(info.state <= MTX_JOB_COMPLETE && other.state <= MTX_JOB_COMPLETE);
    nap();
  }
  CHECK(it < 200000 && info.completed_samples == samples - 3);
  const char *names[3] = {"hdr", "normal", "albedo"};
  for (uint32_t k = 0; k < 3; k++) {
    CHECK(image(a, k, pa, size) == 0 && image(b, k, pb, size) == 0);
    CHECK(memcmp(pa, pb, size) == 0);
    char path[4096];
    if (snprintf(path, sizeof path, "%s/%s.rgba32f", argv[2], names[k]) >= (int)sizeof path) return 2;
    FILE *f = fopen(path, "wb");
    CHECK(f && fwrite(pa, 1, size, f) == size);
    fclose(f);
  }
  /* Size query and a short destination. */
  written = 0;
  OK(mtx_context_read_pixels(a, MTX_IMAGE_HDR, NULL, 0, &written, error, sizeof error));
  CHECK(written == size);
  STATUS(MTX_BUFFER_TOO_SMALL, mtx_context_read_pixels(a, MTX_IMAGE_HDR, pa, size - 1, &written, error, sizeof error));
  CHECK(written == size);
  STATUS(MTX_INVALID_ARGUMENT, mtx_context_read_pixels(a, 3, pa, size, &written, error, sizeof error));

  /* Cancelling after the final batch was submitted still completes. */
  OK(mtx_job_cancel(ja, error, sizeof error));
  OK(mtx_job_info(ja, &info, error, sizeof error));
  CHECK(info.state == MTX_JOB_COMPLETE);
  /* A rejected commit keeps the scene and the job. */
  uint32_t version;
  memcpy(&version, desc, 4);
  const uint32_t bad = 99;
  memcpy(desc, &bad, 4);
  STATUS(MTX_INVALID_ARGUMENT, mtx_context_commit(b, desc, length, error, sizeof error));
  memcpy(desc, &version, 4);
  OK(mtx_job_info(jb, &info, error, sizeof error));
  CHECK(info.state == MTX_JOB_COMPLETE);
  OK(mtx_context_reset(a, error, sizeof error));
  STATUS(MTX_INVALID_HANDLE, mtx_job_info(ja, &info, error, sizeof error));
  OK(mtx_context_commit_with_settings(b, desc, length, &settings, error, sizeof error));
  STATUS(MTX_INVALID_HANDLE, mtx_job_info(jb, &info, error, sizeof error));

  OK(mtx_context_submit(a, 8, &ja, error, sizeof error));
  OK(mtx_job_poll(ja, 64, error, sizeof error));
  OK(mtx_job_cancel(ja, error, sizeof error));
  CHECK(finish(ja, 0, &info) == 0);
  CHECK(info.state == MTX_JOB_COMPLETE && info.submitted_samples == 8 && info.completed_samples == 8);
  OK(mtx_context_render(b, 8, error, sizeof error));
  for (uint32_t k = 0; k < 3; k++) {
    CHECK(image(a, k, pa, size) == 0 && image(b, k, pb, size) == 0);
    CHECK(memcmp(pa, pb, size) == 0);
  }
  OK(mtx_context_render(a, 1, error, sizeof error));
  STATUS(MTX_INVALID_HANDLE, mtx_job_info(ja, &info, error, sizeof error));
  OK(mtx_context_submit(b, 8, &jb, error, sizeof error));
  STATUS(MTX_INVALID_STATE, mtx_context_destroy(b, error, sizeof error));
  OK(mtx_job_cancel(jb, error, sizeof error));
  OK(mtx_context_destroy(b, error, sizeof error));
  STATUS(MTX_INVALID_HANDLE, mtx_job_info(jb, &info, error, sizeof error));
  OK(mtx_context_destroy(a, error, sizeof error));
  STATUS(MTX_INVALID_HANDLE, mtx_context_destroy(a, error, sizeof error));
  free(desc);
  free(pa);
  free(pb);
  printf("PASS %u checks\n", checks);
  return 0;
}
