# Turns one of the reference's driver programs (src/streamer_{2d,3d,cyl}.f90)
# into a module, so that the test harness can call its internal procedures
# and its main loop. Run by oracle/Makefile on the file under the reference
# tree; the output goes to oracle/_ref/src and is never committed.
#
#   awk -v name=2d -f drv2mod.awk streamer_2d.f90 > ref_drv_2d.f90
#
# What changes, and nothing else:
#  - `program streamer_X` / `end program` become `module ref_drv_X` /
#    `end module`; the program's declarations become module variables, so the
#    internal procedures see the same host entities (tree, mg, ...);
#  - `use m_write_silo` and `use m_aD_io` are dropped (Silo is not available),
#    and so are the statements that write output files (aD_write_silo,
#    write_streamer_properties), continuation lines included;
#  - the executable part from `call init_tree(tree)` up to (not including)
#    `call aD_destroy(tree)` becomes the module procedure drv_main, placed
#    first after `contains`. The lines before init_tree read the config
#    files and the transport data; the harness sets those module variables
#    (ST_*, ST_td_tbl, ST_init_cond) itself before it calls drv_main;
#  - two counters for the CPU baseline: at the start of every time step
#    (the line after `do n = 1, ST_steps_amr`) drv_steps is incremented and
#    the step's leaf cells (leaf boxes of levels >= 1 times box_size^dim)
#    are added to drv_cells. They read the tree only.
BEGIN { state = 0; skipping = 0; nbuf = 0; dim = (name == "3d") ? 3 : 2 }

function cont(line) { return line ~ /&[ \t]*$/ }

{
  line = $0
  if (state == 0) {
    if (line ~ /^program[ \t]/) { print "module ref_drv_" name; next }
    if (line ~ /^[ \t]*use m_write_silo/ || line ~ /^[ \t]*use m_a[23]_io/) next
    if (line ~ /^[ \t]*call ST_create_cfg/) { state = 1; next }
    print line
    next
  }
  if (state == 1) {
    if (line ~ /^[ \t]*call init_tree\(tree\)/) { state = 2; buf[++nbuf] = line }
    next
  }
  if (state == 2) {
    if (skipping) { skipping = cont(line); next }
    if (line ~ /^[ \t]*call a[23]_destroy\(tree\)/) { state = 3; next }
    if (line ~ /write_silo|write_streamer_properties/) { skipping = cont(line); next }
    buf[++nbuf] = line
    if (line ~ /^[ \t]*do n = 1, ST_steps_amr[ \t]*$/) {
      buf[++nbuf] = "        drv_steps = drv_steps + 1"
      buf[++nbuf] = "        drv_cells = drv_cells + drv_leaf_cells()"
      ncount++
    }
    next
  }
  if (state == 3) {
    if (line ~ /^contains[ \t]*$/) {
      print "  integer :: drv_steps = 0"
      print "  real(kind(0.0d0)) :: drv_cells = 0"
      print "contains"
      print ""
      print "  subroutine drv_main()"
      for (k = 1; k <= nbuf; k++) print buf[k]
      print "  end subroutine drv_main"
      print ""
      print "  function drv_leaf_cells() result(c)"
      print "    real(kind(0.0d0)) :: c"
      print "    integer :: l"
      print "    c = 0"
      print "    do l = 1, tree%max_lvl"
      print "       c = c + size(tree%lvls(l)%leaves)"
      print "    end do"
      print "    c = c * real(ST_box_size, kind(0.0d0))**" dim
      print "  end function drv_leaf_cells"
      state = 4
    }
    next
  }
  if (line ~ /^end program/) { print "end module ref_drv_" name; next }
  print line
}

END {
  if (state != 4 || ncount != 1) {
    print "drv2mod.awk: " FILENAME " does not have the expected program layout" > "/dev/stderr"
    exit 1
  }
}
