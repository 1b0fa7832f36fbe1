// ExchangeMode.hh -- which per-particle quantity a multi-GPU exchange moves
// (reference ExchangeMode.hh:4-10; the enumerator names and order are kept
// so that driver code naming them compiles unchanged).
#pragma once

enum class ExchangeMode {
  force,
  acceleration,
  torque,
  coordnumber,
  density,
};
