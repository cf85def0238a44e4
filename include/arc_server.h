/*
 * arc_server.h -- device-resident step server over libarcsim.
 *
 * No counterpart in the reference: there every caller owns a whole arc_sim
 * and steps it synchronously (csrc/arc_sim.h:25-56).  Here one server owns
 * the worlds of many clients in one device-resident sim:
 *   - admission by HBM budget: a client asks for n worlds; it is admitted
 *     when the worlds already held plus n fit the budget (bytes per world as
 *     arcx_bytes_per_world reports) and a contiguous free range exists;
 *   - batching: step requests are queued as they arrive; a dispatcher thread
 *     takes the oldest pending request of every client, uploads their host
 *     targets (device targets are read where the client holds them) and
 *     steps all their worlds in one batched launch (arcx_step_batch with the
 *     fall rules and auto-reset, replayed from a hipGraph).  A batch starts
 *     at once when every open client has a request pending; while another
 *     batch runs, shortly before that one is expected to end (so requests
 *     arriving meanwhile go together into the next); with none running, when
 *     requests stop arriving for 100 us, at the latest a batching window
 *     after the oldest;
 *   - latency: each request is timed from its submission to the end of the batch
 *     that stepped it (host steady clock); arcs_stats reports p50 / p99.
 * A client's requests are stepped in submission order, one per batch.
 * Thread safety: every function may be called from any host thread.
 */
#ifndef ARC_SERVER_H
#define ARC_SERVER_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct arc_server arc_server;

/* hbm_budget_bytes: device bytes the server may hold for worlds (the world
 * capacity is budget / bytes per world); dt, substeps, grid, cell as
 * arc_sim_create.  NULL on failure (no GPU, budget below one world). */
arc_server *arcs_create(double hbm_budget_bytes, double dt, int substeps, int grid, double cell);
void arcs_destroy(arc_server *sv);

/* worlds the budget holds, and worlds held by admitted clients */
int arcs_capacity(const arc_server *sv);
int arcs_used(const arc_server *sv);

/* Admission: n_worlds worlds reset with seeds seed0 + k and the given
 * difficulty (< 0: per world uniform in [0, 1) from the engine RNG seeded
 * seed0).  Returns a client id >= 0, or -1 when the budget or the free space
 * cannot hold them or the difficulty exceeds 6 (nothing is changed then).
 * Ids are issued in increasing order and never reused, so an id held after
 * its client was closed names no other client. */
int arcs_open(arc_server *sv, int n_worlds, unsigned int seed0, double difficulty);
/* Releases a client's worlds after its pending requests are stepped. */
int arcs_close(arc_server *sv, int client);
/* first world and world count of a client in the server's sim */
int arcs_client_range(const arc_server *sv, int client, int *start, int *count);

/* Queues one control step of a client: targets = n_worlds x 12 floats
 * (host memory, copied before return).  Returns a ticket > 0, or -1. */
long long arcs_submit(arc_server *sv, int client, const float *targets);
/* Queues one control step of a client whose targets are already in device
 * memory: targets_dev = n_worlds x 12 float32 on the server's device, written
 * completely before the call (the caller synchronises its producer) and left
 * unchanged until arcs_wait on the ticket returns.  The step kernels read them
 * where they lie (no copy; nothing crosses PCIe).
 * Returns a ticket > 0, or -1 (also when targets_dev is not device memory of
 * the server's GPU).  Requests of either kind may be mixed. */
long long arcs_submit_device(arc_server *sv, int client, const float *targets_dev);
/* Blocks until the ticket's step is done; returns its latency in
 * milliseconds (submit to completion), or -1 for an unknown ticket. */
double arcs_wait(arc_server *sv, long long ticket);

/* Runs the sim's autotuner (arcx_autotune) on the worlds held now, between
 * batches; out5 / ms as arcx_autotune.  Call it once the clients are open. */
int arcs_autotune(arc_server *sv, int steps, int *out5, double *ms);

/* Batching window: with no batch running, a batch waits at most this long
 * after its oldest request for requests of the other open clients (less when
 * requests stop arriving for 100 us), and starts at once when every open
 * client has one pending (default 500 us; 0 = never wait). */
int arcs_set_window(arc_server *sv, double microseconds);

/* out[0..9]: requests completed, batches launched, world-steps, mean
 * worlds per batch, p50 / p99 / max latency (ms), elapsed seconds since the
 * last arcs_reset_stats, world-steps per second over it, mean requests per
 * batch. */
int arcs_stats(const arc_server *sv, double *out10);
void arcs_reset_stats(arc_server *sv);

/* root state (13 doubles per world, as arc_sim_observe) of a client's worlds,
 * read after its last completed step */
int arcs_observe(arc_server *sv, int client, double *root);

#ifdef __cplusplus
}
#endif

#endif
