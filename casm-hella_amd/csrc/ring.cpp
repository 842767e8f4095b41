// ring.cpp -- PSRDADA-style block ring in pinned host memory.
//
// casm-hella reads its input from a PSRDADA ring (src/hella/src/dada.cpp:16-88:
// connect, lock, read the text header, then open/close data blocks in order,
// casm:659-670,812-815).  This ring keeps that shape for the streaming
// server: a text header, fixed-size data blocks written in order by one
// producer and read in order by one consumer.  Each block also carries the
// beam it belongs to, the beam's chunk number, flags and the time the
// producer finished writing it (the chunk's arrival time).
//
// The reader may release blocks out of order (the server keeps a block until
// its host-to-device copy has completed); the writer reuses a block only once
// every older block has been released, as a DADA writer waits for the
// reader to clear the oldest block.  Data blocks are allocated with
// hipHostMalloc, so the server copies them to the device without staging; a
// process that sees no GPU (a writer on its own) gets pageable blocks.
#include <chrono>
#include <condition_variable>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "internal.hpp"
#include "pipeline.hpp"

struct hella_ring_s {
  int nblocks = 0;
  size_t block_bytes = 0;
  std::string header;
  uint8_t *data = nullptr;
  bool pinned = false;
  struct Slot {
    int beam = 0, flags = 0;
    long long seq = 0, t_ns = 0;
    size_t bytes = 0;
    bool released = false;
  };
  std::vector<Slot> slot;
  long long w_seq = 0;      // blocks published by the writer
  long long r_seq = 0;      // blocks handed to the reader
  long long clear_seq = 0;  // oldest block not yet released
  bool writing = false;
  bool ended = false;
  bool aborted = false;     // the reader gave up: writers stop waiting
  std::mutex m;
  std::condition_variable cv;
};

namespace hamd {

long long now_ns()
{
  return (long long)std::chrono::duration_cast<std::chrono::nanoseconds>(
             std::chrono::steady_clock::now().time_since_epoch())
      .count();
}

}  // namespace hamd

using namespace hamd;

namespace {

hella_status ring_fail(hella_status s, const char *msg)
{
  set_last_error(msg);
  return s;
}

template <class Pred>
bool wait_for(hella_ring r, std::unique_lock<std::mutex> &lk, int timeout_ms, Pred pred)
{
  if (timeout_ms < 0) {
    r->cv.wait(lk, pred);
    return true;
  }
  return r->cv.wait_for(lk, std::chrono::milliseconds(timeout_ms), pred);
}

// memcpy split over a fixed set of worker threads (the caller copies the
// first slice); workers live as long as the object
class ParallelCopy {
 public:
  explicit ParallelCopy(int nthreads) : n_(std::max(1, nthreads))
  {
    for (int i = 1; i < n_; ++i) th_.emplace_back([this, i] { work(i); });
  }
  ~ParallelCopy()
  {
    {
      std::lock_guard<std::mutex> lk(m_);
      stop_ = true;
    }
    cv_.notify_all();
    for (auto &t : th_) t.join();
  }
  void copy(void *dst, const void *src, size_t bytes)
  {
    if (n_ == 1) {
      std::memcpy(dst, src, bytes);
      return;
    }
    {
      std::lock_guard<std::mutex> lk(m_);
      dst_ = static_cast<uint8_t *>(dst);
      src_ = static_cast<const uint8_t *>(src);
      bytes_ = bytes;
      pending_ = n_ - 1;
      gen_ += 1;
    }
    cv_.notify_all();
    slice(0);
    std::unique_lock<std::mutex> lk(m_);
    done_cv_.wait(lk, [&] { return pending_ == 0; });
  }

 private:
  void slice(int i)
  {
    const size_t per = (bytes_ / (size_t)n_ + 63) & ~(size_t)63;
    const size_t a = std::min(bytes_, per * (size_t)i), b = std::min(bytes_, a + per);
    if (b > a) std::memcpy(dst_ + a, src_ + a, b - a);
  }
  void work(int i)
  {
    long long seen = 0;
    for (;;) {
      {
        std::unique_lock<std::mutex> lk(m_);
        cv_.wait(lk, [&] { return stop_ || gen_ != seen; });
        if (stop_) return;
        seen = gen_;
      }
      slice(i);
      {
        std::lock_guard<std::mutex> lk(m_);
        pending_ -= 1;
      }
      done_cv_.notify_one();
    }
  }
  int n_;
  std::vector<std::thread> th_;
  std::mutex m_;
  std::condition_variable cv_, done_cv_;
  bool stop_ = false;
  long long gen_ = 0;
  int pending_ = 0;
  uint8_t *dst_ = nullptr;
  const uint8_t *src_ = nullptr;
  size_t bytes_ = 0;
};

}  // namespace

extern "C" {

hella_status hella_ring_create(int nblocks, size_t block_bytes, const char *header, hella_ring *out)
{
  if (!out || nblocks <= 0 || block_bytes == 0) return ring_fail(HELLA_ERR_INVALID, "bad ring size");
  *out = nullptr;
  auto *r = new (std::nothrow) hella_ring_s();
  if (!r) return ring_fail(HELLA_ERR_ALLOC, "ring allocation failed");
  r->nblocks = nblocks;
  r->block_bytes = block_bytes;
  r->header = header ? header : "";
  r->slot.resize((size_t)nblocks);
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) == hipSuccess && ndev > 0) {
    if (hipHostMalloc((void **)&r->data, (size_t)nblocks * block_bytes, hipHostMallocDefault) !=
        hipSuccess) {
      delete r;
      return ring_fail(HELLA_ERR_ALLOC, "pinned ring allocation failed");
    }
    r->pinned = true;
  } else {
    r->data = static_cast<uint8_t *>(std::malloc((size_t)nblocks * block_bytes));
    if (!r->data) {
      delete r;
      return ring_fail(HELLA_ERR_ALLOC, "ring allocation failed");
    }
  }
  *out = r;
  return HELLA_OK;
}

void hella_ring_destroy(hella_ring r)
{
  if (!r) return;
  if (r->data) {
    if (r->pinned)
      (void)hipHostFree(r->data);
    else
      std::free(r->data);
  }
  delete r;
}

const char *hella_ring_header(hella_ring r) { return r ? r->header.c_str() : nullptr; }

hella_status hella_ring_open_write(hella_ring r, int timeout_ms, void **block)
{
  if (!r || !block) return ring_fail(HELLA_ERR_INVALID, "bad argument");
  std::unique_lock<std::mutex> lk(r->m);
  if (r->writing) return ring_fail(HELLA_ERR_STATE, "a block is already open for writing");
  if (r->ended) return ring_fail(HELLA_ERR_STATE, "the ring has ended");
  if (!wait_for(r, lk, timeout_ms,
                [&] { return r->aborted || r->w_seq - r->clear_seq < r->nblocks; }))
    return ring_fail(HELLA_ERR_CAPACITY, "no free block (timeout)");
  if (r->aborted) return ring_fail(HELLA_ERR_STATE, "the reader aborted the ring");
  r->writing = true;
  *block = r->data + (size_t)(r->w_seq % r->nblocks) * r->block_bytes;
  return HELLA_OK;
}

hella_status hella_ring_close_write(hella_ring r, int beam, long long seq, int flags, size_t bytes)
{
  if (!r) return ring_fail(HELLA_ERR_INVALID, "ring is NULL");
  {
    std::lock_guard<std::mutex> lk(r->m);
    if (!r->writing) return ring_fail(HELLA_ERR_STATE, "no block open for writing");
    if (bytes > r->block_bytes) return ring_fail(HELLA_ERR_INVALID, "block overflow");
    auto &s = r->slot[(size_t)(r->w_seq % r->nblocks)];
    s.beam = beam;
    s.seq = seq;
    s.flags = flags;
    s.bytes = bytes;
    s.released = false;
    s.t_ns = now_ns();
    r->w_seq += 1;
    r->writing = false;
  }
  r->cv.notify_all();
  return HELLA_OK;
}

hella_status hella_ring_end(hella_ring r)
{
  if (!r) return ring_fail(HELLA_ERR_INVALID, "ring is NULL");
  {
    std::lock_guard<std::mutex> lk(r->m);
    r->ended = true;
  }
  r->cv.notify_all();
  return HELLA_OK;
}

hella_status hella_ring_abort(hella_ring r)
{
  if (!r) return ring_fail(HELLA_ERR_INVALID, "ring is NULL");
  {
    std::lock_guard<std::mutex> lk(r->m);
    r->aborted = true;
    r->ended = true;
  }
  r->cv.notify_all();
  return HELLA_OK;
}

hella_status hella_ring_open_read(hella_ring r, int timeout_ms, hella_ring_block *info)
{
  if (!r || !info) return ring_fail(HELLA_ERR_INVALID, "bad argument");
  std::unique_lock<std::mutex> lk(r->m);
  std::memset(info, 0, sizeof(*info));
  info->index = -1;
  wait_for(r, lk, timeout_ms, [&] { return r->r_seq < r->w_seq || r->ended; });
  if (r->r_seq < r->w_seq) {
    const int idx = (int)(r->r_seq % r->nblocks);
    const auto &s = r->slot[(size_t)idx];
    info->index = idx;
    info->beam = s.beam;
    info->seq = s.seq;
    info->flags = s.flags;
    info->bytes = s.bytes;
    info->t_arrival_ns = s.t_ns;
    info->data = r->data + (size_t)idx * r->block_bytes;
    r->r_seq += 1;
    return HELLA_OK;
  }
  if (r->ended) info->flags = HELLA_RING_END;
  return HELLA_OK;
}

hella_status hella_ring_release(hella_ring r, int index)
{
  if (!r || index < 0 || index >= r->nblocks) return ring_fail(HELLA_ERR_INVALID, "bad block index");
  {
    std::lock_guard<std::mutex> lk(r->m);
    r->slot[(size_t)index].released = true;
    while (r->clear_seq < r->r_seq && r->slot[(size_t)(r->clear_seq % r->nblocks)].released)
      r->clear_seq += 1;
  }
  r->cv.notify_all();
  return HELLA_OK;
}

hella_status hella_ring_replay(hella_ring r, const hella_replay_event *ev, int n,
                               const void *const *pool, int npool, size_t chunk_bytes,
                               double time_scale, long long *max_late_ns)
{
  if (!r || (n > 0 && (!ev || !pool)) || chunk_bytes > r->block_bytes || time_scale < 0)
    return ring_fail(HELLA_ERR_INVALID, "bad argument");
  // copies of large chunks are split over a few threads: one host thread
  // cannot fill blocks as fast as the server drains them
  ParallelCopy pc(chunk_bytes >= (4u << 20) ? 8 : 1);
  const long long t0 = now_ns();
  long long late = 0;
  hella_status st = HELLA_OK;
  for (int i = 0; i < n && st == HELLA_OK; ++i) {
    const hella_replay_event &e = ev[i];
    if (e.pool < 0 || e.pool >= npool) {
      st = ring_fail(HELLA_ERR_INVALID, "pool index out of range");
      break;
    }
    const long long due = t0 + (long long)(e.t * time_scale * 1e9);
    const long long wait = due - now_ns();
    if (wait > 0) std::this_thread::sleep_for(std::chrono::nanoseconds(wait));
    void *blk = nullptr;
    st = hella_ring_open_write(r, -1, &blk);
    if (st != HELLA_OK) break;
    pc.copy(blk, pool[e.pool], chunk_bytes);
    late = std::max(late, now_ns() - due);
    st = hella_ring_close_write(r, e.beam, e.seq, e.flags, chunk_bytes);
  }
  if (max_late_ns) *max_late_ns = late;
  // the reader always sees the end of the stream, also after an error
  const hella_status se = hella_ring_end(r);
  return st != HELLA_OK ? st : se;
}

}  // extern "C"
