# The reference's configuration (apps/sgemm/tests.mk:1) ...
TESTS += $(call test-name,512,2)
# ... and the MI355X workloads: precision and input kinds are optional fields.
TESTS += $(call test-name,512,2)__INPUT_uniform
TESTS += $(call test-name,512,2)__PREC_bf16x3
TESTS += $(call test-name,8192,10)__INPUT_uniform
TESTS += $(call test-name,8192,1)__PREC_bf16x3__INPUT_loguniform
# ... and the grouped path over the same slices: size heuristic, a forced tile
# shape and split-K, and the autotuner's tuning file.
TESTS += $(call test-name,512,2)__MODE_grouped
TESTS += $(call test-name,640,3)__MODE_grouped__PREC_bf16x3__TILE_0__SPLIT_2
TESTS += $(call test-name,512,2)__MODE_grouped__TUNING_mi355x
