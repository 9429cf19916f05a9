// group_commit.h -- coalescing of concurrent single-query calls into batches.
//
// The reference serves PySearcher::search by running Index::search once per
// query from a pool of threads (ParallelFor, pyapi_wrapper/py_api_wrapper.cpp:
// 111-152), each call re-entrant over read-only index data
// (puck/puck_data_pool.h:75-99).  On the GPU one call per query would launch a
// whole pipeline for one query, so concurrent callers are grouped: a call
// enqueues its item; when no batch is being run, the caller becomes the leader,
// takes every queued item (up to max_batch) and runs them as one batch; callers
// arriving meanwhile queue up and form the next batch.  A lone caller runs at
// once (no timer, no added latency); under load the batch grows to whatever
// arrived while the previous one ran.
//
// Each submit() returns after its own item was processed.  A leader's batch
// holds its own item first, then the oldest queued ones; after the batch it
// hands leadership to a waiting caller whose item is still queued and
// returns, so no caller serves batches for others indefinitely.
#pragma once
#include <algorithm>
#include <condition_variable>
#include <cstddef>
#include <cstdint>
#include <deque>
#include <mutex>
#include <vector>

namespace puck {

template <typename Item>
class GroupCommit {
public:
    explicit GroupCommit(size_t max_batch) : _max(max_batch < 1 ? 1 : max_batch) {}

    // Runs fn(items, n) on a batch that includes *item and returns when that
    // batch has finished.  fn runs on one thread at a time.
    template <typename Fn>
    void submit(Item* item, Fn&& fn) {
        Node node{item, false};
        std::unique_lock<std::mutex> lk(_mu);
        _queue.push_back(&node);
        _cv.wait(lk, [&] { return node.done || !_busy; });
        if (node.done) return;
        // leader: its own node (still queued), then the oldest others
        _busy = true;
        std::vector<Node*> batch{&node};
        std::vector<Item*> items;
        _queue.erase(std::find(_queue.begin(), _queue.end(), &node));
        const size_t more = std::min(_max - 1, _queue.size());
        batch.insert(batch.end(), _queue.begin(), _queue.begin() + (std::ptrdiff_t)more);
        _queue.erase(_queue.begin(), _queue.begin() + (std::ptrdiff_t)more);
        const size_t n = batch.size();
        lk.unlock();
        items.reserve(n);
        for (Node* b : batch) items.push_back(b->item);
        fn(items.data(), items.size());
        lk.lock();
        for (Node* b : batch) b->done = true;
        ++_batches;
        _items += n;
        _max_seen = std::max<uint64_t>(_max_seen, n);
        _busy = false;
        lk.unlock();
        _cv.notify_all();
    }

    uint64_t batches() const {
        std::lock_guard<std::mutex> g(_mu);
        return _batches;
    }
    uint64_t items() const {
        std::lock_guard<std::mutex> g(_mu);
        return _items;
    }
    uint64_t largest_batch() const {
        std::lock_guard<std::mutex> g(_mu);
        return _max_seen;
    }

private:
    struct Node {
        Item* item;
        bool done;
    };
    const size_t _max;
    mutable std::mutex _mu;
    std::condition_variable _cv;
    std::deque<Node*> _queue;
    bool _busy = false;
    uint64_t _batches = 0, _items = 0, _max_seen = 0;
};

}  // namespace puck
