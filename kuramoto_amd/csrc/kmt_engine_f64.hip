// FP64 instantiation of the single-network engine (its own translation
// unit, so the per-precision kernel sets compile in parallel).
#include "kmt_engine.hpp"

namespace kmt {
EngineBase *new_engine_f64() { return new Engine<double>(); }
}  // namespace kmt
