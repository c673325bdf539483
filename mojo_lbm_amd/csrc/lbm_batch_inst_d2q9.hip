// Batched-case launchers of the Q=9 velocity set, all storages.
#include "lbm_batch_impl.hpp"

namespace lbm {
template struct BatchLaunch<Lattice<9>, StoreF32>;
template struct BatchLaunch<Lattice<9>, StoreF16C>;
template struct BatchLaunch<Lattice<9>, StoreF64>;
}  // namespace lbm
