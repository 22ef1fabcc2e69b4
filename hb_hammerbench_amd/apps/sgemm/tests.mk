# The reference's configuration (apps/sgemm/tests.mk:1) ...
TESTS += $(call test-name,512,2)
# ... and the MI355X workloads: precision and input kinds are optional fields.
TESTS += $(call test-name,512,2)__INPUT_uniform
TESTS += $(call test-name,512,2)__PREC_bf16x3
TESTS += $(call test-name,8192,10)__INPUT_uniform
TESTS += $(call test-name,8192,1)__PREC_bf16x3__INPUT_loguniform
