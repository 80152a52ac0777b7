"""Checker package: the FP64 CPU restatement of the reference treecode and the
shim over the reference itself.  Test infrastructure only -- the product path
(bit_amd/) never imports it."""
