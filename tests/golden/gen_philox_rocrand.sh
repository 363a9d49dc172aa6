#!/bin/sh
# Regenerates philox_rocrand.txt from rocRAND's Philox source header (host build).
set -e
cd "$(dirname "$0")"
hipcc -x hip --offload-arch=gfx950 -O1 -o /tmp/gen_philox_rocrand gen_philox_rocrand.cpp
/tmp/gen_philox_rocrand > philox_rocrand.txt
