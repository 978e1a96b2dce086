// Host side of the nearest-neighbour backend (tapkee's find_neighbors,
// include/tapkee/neighbors/neighbors.hpp:119-216): the connectivity test of
// a neighbour graph and a way for front ends to log through the engine's
// logger.  The neighbours themselves come from the GPU kNN (tka_knn).
#include <string>
#include <vector>

#include "common.hpp"

using namespace tka;

extern "C" {

// include/tapkee/neighbors/connected.hpp:19-61: depth-first search from
// object 0 along the directed neighbour lists; connected when it reaches
// every object.
TKA_API tka_status tka_neighbors_connected(const int32_t* idx, int32_t N, int32_t k, int32_t* connected)
{
    return guard([&] {
        require(N >= 1 && k >= 0 && (idx || k == 0) && connected, TKA_E_WRONG_PARAMETER, "bad arguments");
        std::vector<char> seen((size_t)N, 0);
        std::vector<int32_t> stack;
        stack.reserve((size_t)N);
        stack.push_back(0);
        int32_t reached = 0;
        while (!stack.empty() && reached < N) {
            const int32_t v = stack.back();
            stack.pop_back();
            if (seen[(size_t)v]) continue;
            seen[(size_t)v] = 1;
            ++reached;
            const int32_t* nb = idx + (size_t)v * k;
            for (int32_t j = 0; j < k; j++) {
                const int32_t u = nb[j];
                require(u >= 0 && u < N, TKA_E_WRONG_PARAMETER, "neighbour index out of range");
                if (!seen[(size_t)u]) stack.push_back(u);
            }
        }
        *connected = reached == N ? 1 : 0;
    });
}

TKA_API void tka_log(int32_t level, const char* message)
{
    if (level < TKA_LOG_INFO || level > TKA_LOG_BENCHMARK || message == nullptr) return;
    log_message(static_cast<tka_log_level>(level), message);
}

}  // extern "C"
