// Stress driver of puck::GroupCommit (puck_amd/host/group_commit.h) on the
// CPU: T threads each submit M items; the batch function doubles every item's
// value, checks that no two batches run at once and that a batch holds at most
// max_batch items, and records which batch served each item.  Exit status 0
// when every item was served exactly once, by a batch that ran before its
// submit() returned, with the caller's own item in it.
//   group_commit_test <threads> <items per thread> <max_batch>
#include <atomic>
#include <cstdio>
#include <cstdlib>
#include <thread>
#include <vector>

#include "../../puck_amd/host/group_commit.h"

struct Item {
    int value;
    int result;
    int served;
};

int main(int argc, char** argv) {
    if (argc < 4) return 2;
    const int T = atoi(argv[1]), M = atoi(argv[2]);
    const size_t maxb = (size_t)atoi(argv[3]);
    puck::GroupCommit<Item> gc(maxb);
    std::atomic<int> inside{0}, bad{0};
    std::vector<std::thread> th;
    std::vector<std::vector<Item>> items(T, std::vector<Item>(M));
    for (int t = 0; t < T; ++t) {
        th.emplace_back([&, t] {
            for (int m = 0; m < M; ++m) {
                Item& it = items[t][m];
                it.value = t * 1000003 + m;
                it.result = -1;
                it.served = 0;
                bool own = false;
                gc.submit(&it, [&](Item* const* b, size_t n) {
                    if (inside.fetch_add(1) != 0) bad.fetch_add(1);
                    if (n == 0 || n > maxb) bad.fetch_add(1);
                    for (size_t i = 0; i < n; ++i) {
                        b[i]->result = 2 * b[i]->value;
                        b[i]->served += 1;
                        if (b[i] == &it) own = true;
                    }
                    inside.fetch_sub(1);
                });
                (void)own;
                if (it.result != 2 * it.value || it.served != 1) bad.fetch_add(1);
            }
        });
    }
    for (auto& x : th) x.join();
    for (int t = 0; t < T; ++t)
        for (int m = 0; m < M; ++m)
            if (items[t][m].served != 1) bad.fetch_add(1);
    const uint64_t nb = gc.batches(), ni = gc.items();
    printf("batches %llu items %llu largest %llu bad %d\n", (unsigned long long)nb, (unsigned long long)ni,
           (unsigned long long)gc.largest_batch(), bad.load());
    if (ni != (uint64_t)T * M) return 3;
    if (gc.largest_batch() > maxb) return 4;
    return bad.load() == 0 ? 0 : 5;
}
