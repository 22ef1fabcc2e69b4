# Per-test Makefile (copied into each generated directory), the counterpart
# of apps/sgemm/template.mk:1-78: parameters become launcher arguments.
include parameters.mk
include app_path.mk

SGEMM ?= $(APP_PATH)/sgemm
ARGS  := --N $(N) --NITER $(NITER) --precision $(PREC) --input $(INPUT)
ifeq ($(MODE),grouped)
ARGS  += --mode grouped
ifneq ($(TILE),auto)
ARGS  += --tile $(TILE) --split $(SPLIT)
else ifneq ($(TUNING),none)
ARGS  += --tuning $(APP_PATH)/../../tuning/$(TUNING)_grouped.txt
endif
endif

.PHONY: help
help:
	@echo "targets: exec.log profile.log ($(test-name))"

exec.log: $(SGEMM)
	$(SGEMM) $(ARGS) --reps 3 2>&1 | tee $@

profile.log: $(SGEMM)
	cd /tmp && rocprofv3 --kernel-trace --stats -f csv -d $(CURDIR)/profile -o run -- $(SGEMM) $(ARGS) --reps 3 2>&1 | tee $(CURDIR)/$@
