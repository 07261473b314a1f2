# Compiler flags shared by the library build (Makefile) and the device
# builds of its test harnesses (tests/hosttrace/Makefile), so that a harness
# compiles the traversal code exactly as the library does.
QBT_CSRC := $(abspath $(dir $(lastword $(MAKEFILE_LIST))))
QBT_ROOT := $(abspath $(QBT_CSRC)/../..)
HIPCC    ?= /opt/rocm/bin/hipcc
CC       ?= gcc
ARCH     ?= gfx950

CFLAGS   := -O2 -std=gnu11 -fPIC -ffp-contract=off -Wall -Wextra -Wno-unused-parameter -pthread -I$(QBT_ROOT)/include
HIPFLAGS := --offload-arch=$(ARCH) -O3 -std=c++17 -fPIC -ffp-contract=off \
            -fhip-fp32-correctly-rounded-divide-sqrt -fno-gpu-flush-denormals-to-zero \
            -Xclang -target-feature -Xclang -packed-fp32-ops \
            -Wall -Wno-unused-parameter -I$(QBT_ROOT)/include -I$(QBT_CSRC)/device
