// Fused NTT kernels with 8192-word tiles for the two-pass sizes.
#include "ntt_fused.h"

namespace vk {

static const std::array<FusedEntry, kFusedEntries> g_table =
	make_fused_table<13>(std::make_integer_sequence<int, kFusedEntries>());

const FusedEntry *fused_table_kt13() {
	return g_table.data();
}

}  // namespace vk
