APP_PATH := $(abspath $(dir $(lastword $(MAKEFILE_LIST)))..)
