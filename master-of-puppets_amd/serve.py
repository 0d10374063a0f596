"""Serving form: a request loop over resident scenes with continuous batching.

Resident state: every scene's meshes, IBL tables and shadow map stay in HBM
(the shadow map is rendered once and reused: it depends on the scene and its
lights, not on the camera).  Requests arrive as a Poisson stream; each picks a
scene, an orbit delta of the scene camera and a resolution class.

Admission is by HBM budget.  A share of the budget is cut into framebuffer
slots per resolution class (56 B/px on the device); the rest bounds the
scratch a batch allocates (80 B per transformed vertex, 80 B per triangle for
the triangle records, big-triangle queues and the clip store, plus a fixed
checkpoint store).  A request is admitted into the next batch when a slot of
its class is free and the batch's scratch stays within that bound; it waits
in a bounded queue otherwise, and is rejected when the queue is full.  Continuous batching: while up to `inflight` batches run on the
device, newly arrived requests are recorded into free slots and launched as
the next batch; finished batches are read back (RGBA8 to host) and their
slots freed (each result's RGBA8 copy is queued on the copy stream when its
batch is launched, so copies overlap the next batches' kernels).  Latency =
arrival to the result on the host.  The plain mode serves the
same stream one request per launch, synchronously.
"""
import collections
import math
import time

import numpy as np

from . import driver

RES_CLASSES = [((640, 480), 0.25), ((1280, 720), 0.50), ((1920, 1080), 0.25)]
# device bytes of one framebuffer pixel: colour x2 (8), HDR (16), depth (4),
# id (4), visibility (8), edge values (16)
BYTES_PER_PX = 56
# device scratch of a batch, grown to the largest batch and kept: per
# transformed vertex (clip and screen positions, three attribute float4) and
# per triangle (record 16, big-triangle queue, offsets and checkpoint index
# 12, clip store 768 B per clip slot at one slot per 16 triangles), plus the
# checkpoint store's floor
SCRATCH_PER_VERTEX = 80
SCRATCH_PER_TRIANGLE = 80
SCRATCH_FIXED = 64 << 20
SLOT_SHARE = 0.25


def make_requests(rate, duration, n_scenes, seed=0x4D4F5005):
    rng = np.random.default_rng(seed)
    n = int(rate * duration * 1.2) + 16
    t = np.cumsum(rng.exponential(1.0 / rate, n))
    t = t[t < duration]
    k = len(t)
    scene = rng.integers(0, n_scenes, k)
    yaw = rng.uniform(-5.0, 5.0, k)
    pitch = rng.uniform(-5.0, 5.0, k)
    cls = rng.choice(len(RES_CLASSES), k, p=[p for _, p in RES_CLASSES])
    return [(float(t[i]), int(scene[i]), float(yaw[i]), float(pitch[i]), int(cls[i]))
            for i in range(k)]


class Server:
    def __init__(self, device, scenes, budget_bytes=8 << 30, max_batch=32, inflight=2,
                 max_slots_per_class=64, max_batch_tris=(1 << 27) - 1):
        self.d = device
        self.scenes = scenes  # list of (SceneDriver, Scene)
        self.max_batch = max_batch
        # a batch stays under 2^27 triangles: the fused order-free path keys
        # fragments by (triangle << 4 | fan index) in 32 bits
        self.max_batch_tris = max_batch_tris
        self.tris = [sc.triangle_count for _, sc in scenes]
        self.scratch = [SCRATCH_PER_VERTEX * sum(len(m.vertices) for m in sc.meshes) +
                        SCRATCH_PER_TRIANGLE * sc.triangle_count for _, sc in scenes]
        self.inflight = inflight
        self.budget = budget_bytes
        self.free = []
        self.slots = []
        slot_bytes = 0
        for (w, h), share in RES_CLASSES:
            n = int(budget_bytes * SLOT_SHARE * share // (w * h * BYTES_PER_PX))
            n = max(1, min(n, max_slots_per_class))
            fbs = [device.framebuffer(w, h) for _ in range(n)]
            slot_bytes += n * w * h * BYTES_PER_PX
            self.slots.append(fbs)
            self.free.append(list(fbs))
        self.slot_bytes = slot_bytes
        self.scratch_budget = budget_bytes - slot_bytes - SCRATCH_FIXED
        if self.scratch_budget < max(self.scratch):
            raise ValueError("HBM budget below one request's scratch")
        self.queue_cap = 4 * sum(len(s) for s in self.slots)

    def admission(self):
        """How the HBM budget is divided."""
        return {"budget_bytes": self.budget, "slot_bytes": self.slot_bytes,
                "slots_per_class": {f"{w}x{h}": len(f) for ((w, h), _), f in
                                    zip(RES_CLASSES, self.slots)},
                "scratch_budget_bytes": self.scratch_budget,
                "scratch_bytes_per_request_max": max(self.scratch),
                "max_batch_triangles": self.max_batch_tris}

    def close(self):
        for fbs in self.slots:
            for fb in fbs:
                fb.close()

    def _camera(self, req):
        _, si, yaw, pitch, _ = req
        return driver.camera_of(self.scenes[si][1], yaw, pitch)

    def run(self, requests, plain=False, time_limit=None, keep=()):
        """Serve `requests` ((arrival s, scene, yaw, pitch, class) tuples).
        keep: request indices whose results are kept (in the result's
        "outputs", index -> dict of RGBA8 colour, depth, object ID and HDR
        arrays) to compare with the same request rendered alone or by the
        reference."""
        lat = []
        rejected = 0
        max_queue = 0
        queue = collections.deque()
        inflight = collections.deque()  # (t_launch, [(req, fb)])
        i = 0
        n = len(requests)
        t0 = time.perf_counter()
        done_count = 0
        last_done = t0
        batches = 0
        batch_sizes = []
        keep = set(keep)
        outputs = {}
        index = {id(r): k for k, r in enumerate(requests)}
        while done_count + rejected < n:
            now = time.perf_counter() - t0
            if time_limit and now > time_limit:
                break
            while i < n and requests[i][0] <= now:
                if len(queue) >= self.queue_cap:
                    rejected += 1
                else:
                    queue.append(requests[i])
                i += 1
            max_queue = max(max_queue, len(queue))
            # completions
            progressed = False
            while inflight and all(fb.readback_done() for _, fb in inflight[0][1]):
                _, items = inflight.popleft()
                for req, fb in items:
                    fb.read_color_raw()
                    k = index.get(id(req))
                    if k in keep:
                        outputs[k] = fb.read(color=True, depth=True, ids=True, hdr=True)
                    tdone = time.perf_counter() - t0
                    lat.append(tdone - req[0])
                    self.free[req[4]].append(fb)
                    done_count += 1
                    last_done = tdone
                progressed = True
            # launch
            if queue and len(inflight) < self.inflight:
                batch = []
                skipped = collections.deque()
                limit = 1 if plain else self.max_batch
                tris = 0
                scratch = 0
                while queue and len(batch) < limit:
                    req = queue.popleft()
                    if batch and (tris + self.tris[req[1]] > self.max_batch_tris or
                                  scratch + self.scratch[req[1]] > self.scratch_budget):
                        queue.appendleft(req)
                        break
                    if self.free[req[4]]:
                        tris += self.tris[req[1]]
                        scratch += self.scratch[req[1]]
                        fb = self.free[req[4]].pop()
                        batch.append((req, fb))
                    else:
                        skipped.append(req)
                        if plain:
                            break
                queue.extendleft(reversed(skipped))
                if batch:
                    for req, fb in batch:
                        self.scenes[req[1]][0].record(fb, self._camera(req))
                    self.d.execute([fb for _, fb in batch])
                    # each RGBA8 result is copied to the host on the copy
                    # stream as soon as its batch ends, behind the next
                    # batches' kernels
                    for _, fb in batch:
                        fb.readback_async(color=True, depth=False, ids=False)
                    inflight.append((time.perf_counter() - t0, batch))
                    batches += 1
                    batch_sizes.append(len(batch))
                    progressed = True
                    if plain:
                        # one request per launch, synchronously: wait for it
                        # here; the completion step above reads it back
                        self.d.sync()
                    continue
            if not progressed:
                if i < n and not queue and not inflight:
                    wait = requests[i][0] - (time.perf_counter() - t0)
                    if wait > 0:
                        time.sleep(min(wait, 0.002))
                else:
                    time.sleep(20e-6)
        # a run cut off by its time limit leaves batches on the device: wait
        # for them and give their slots back (they are not counted)
        if inflight:
            self.d.sync()
            for _, items in inflight:
                for req, fb in items:
                    self.free[req[4]].append(fb)
        unfinished = n - done_count - rejected
        lat = np.array(lat) * 1e3
        span = max(last_done, 1e-9)
        return {
            "completed": done_count,
            "rejected": rejected,
            "unfinished": unfinished,
            "max_queue": max_queue,
            "batches": batches,
            "mean_batch": float(np.mean(batch_sizes)) if batch_sizes else 0.0,
            "p50_ms": float(np.percentile(lat, 50)) if len(lat) else None,
            "p99_ms": float(np.percentile(lat, 99)) if len(lat) else None,
            "sustained_fps": done_count / span,
            "outputs": outputs,
        }
