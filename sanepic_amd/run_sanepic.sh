#!/bin/bash
# Drop-in for patanch/SANEPIC's scripts/run_sanepic.sh: the same parameters,
# the same flags (run_sanepic.sh:55-73) and the same input tree, driving the
# MI355X executable.  mpirun is optional: the program uses one GPU and ranks
# other than 0 exit at once, so `mpirun -n ${nprocs}` may be kept.

#################################
##### Setting the arguments #####
#################################

### Number of MPI slots (accepted for compatibility; 1 is enough)
nprocs=${nprocs:-1}

### Base directory
base_dir=${base_dir:-"$PWD/tests"}

### Output location
out_dir="${base_dir}/out_dir/map_"  # For -O

### TOD location
tod_data_dir="${base_dir}/tod_list/"  # For -F
dataSuffixe="output"  # For -B

### Pointings location
pointing_dir="${base_dir}/pointings_list/pointings_"  # For -Z

### ispf location
ispf_dir="${base_dir}/bolo_info"
fknee=${fknee:-0.05}  # For -k
fsamp2=${fsamp2:-9.5}  # For -k
indexN=${indexN:-1.0}  # For -k

### Other params
NSIDE=${NSIDE:-256}
det_num=${det_num:-1}  # For -d
det_name=${det_name:-"LFT6_27_140b"}  # For -C
it="_N${NSIDE}_.bin"  # For -e
pol=1  # For -p

### Bolometer file
bolofile="${base_dir}/bolo_info/bolometer_info.txt"  # For -X

### Segments info
segment_dir="${base_dir}/segments" # For -f
nn=${nn:-12}  # For -f and -n
seglength=${seglength:-256}  # For -u, -f and -l

### HWP params
hwp_status=${hwp_status:-0}  # For -h
OmegaHWPsamp=${OmegaHWPsamp:-1.123456789123}  # For -w

HERE="$(cd "$(dirname "${BASH_SOURCE[0]}")" && pwd)"
EXE="${HERE}/bin/sanepic"

###########################
##### Calling SANEPIC #####
###########################

launcher=()
if [ "${nprocs}" -gt 1 ] && command -v mpirun >/dev/null 2>&1; then
  launcher=(mpirun -n "${nprocs}")
fi

"${launcher[@]}" "${EXE}" \
-F $tod_data_dir \
-B $dataSuffixe \
-f ${segment_dir}/segment_list_${nn}_${seglength}.bi \
-d $det_num \
-C $det_name \
-e $dataSuffixe$it \
-h $hwp_status \
-w $OmegaHWPsamp \
-k ${ispf_dir}/iSpf${fknee}_${fsamp2}_index${indexN}__ \
-u $seglength \
-O $out_dir \
-p $pol \
-l $seglength \
-n $nn \
-N $NSIDE \
-X $bolofile \
-Z $pointing_dir \
-i 0
