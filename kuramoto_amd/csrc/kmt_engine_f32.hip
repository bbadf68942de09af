// FP32 instantiation of the single-network engine (its own translation
// unit, so the per-precision kernel sets compile in parallel).
#include "kmt_engine.hpp"

namespace kmt {
EngineBase *new_engine_f32() { return new Engine<float>(); }
}  // namespace kmt
