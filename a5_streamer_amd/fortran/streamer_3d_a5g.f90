!> The 3D streamer model of src/streamer_3d.f90 driven through m_a5g: the
!> same main loop and step order (src/streamer_3d.f90:57-143), with the box
!> procedures and the multigrid on the GPU and the initial conditions,
!> refinement criterion and time-step rule on the host.
!>
!> Configuration: the 3D air streamer of the production run (a 64 mm cube
!> of one base box of 8^3 cells, as the reference's init_tree builds it,
!> 2.0 MV/m, Gaussian seed of 1e20 m^-3 and width 0.4 mm at (32, 32, 56) mm
!> over 1e15 m^-3, no dielectric, mobility 0.03, diffusion 0.2,
!> alpha(E) = 4.3e5 exp(-2e7/E), eta = 0 given as 500 rows on 0..3e7 V/m and
!> resampled to 1000). Arguments: lvls_max n_blocks [applied_fld [domain_len]].
!> After the initial refinement and after each block of steps_amr steps it
!> prints one line: block, time, dt, max_id, n_add, n_rm, max n_e, min and
!> max |E|, the reals as the hexadecimal bits of the double.
program streamer_3d_a5g
  use iso_c_binding, only: c_int64_t
  use m_a5g
  implicit none

  integer, parameter :: dp = kind(0.0d0)
  real(dp), parameter :: UC_eps0 = 8.8541878176e-12_dp       ! fosito/m_units_constants.f90:34
  real(dp), parameter :: UC_elem_charge = 1.6022e-19_dp      ! fosito/m_units_constants.f90:35

  ! variables (src/m_streamer.f90:16-43)
  integer, parameter :: i_elec = 1, i_pion = 2, i_elec_old = 3, i_pion_old = 4
  integer, parameter :: i_phi = 5, i_fld = 6, i_rhs = 7, i_pho = 8, i_eps = 9
  integer, parameter :: n_var_cell = 9, n_var_face = 2, f_elec = 1, f_fld = 2
  integer, parameter :: i_mobility = 1, i_diffusion = 2, i_alpha = 3, i_eta = 4

  ! configuration
  real(dp) :: domain_len = 64e-3_dp, applied_fld = 2.0e6_dp, bg_dens = 1e15_dp
  real(dp) :: seed_dens = 1e20_dp, seed_width = 0.4e-3_dp
  real(dp) :: seed_r0(3) = [32e-3_dp, 32e-3_dp, 56e-3_dp]
  real(dp) :: seed_r1(3) = [32e-3_dp, 32e-3_dp, 56e-3_dp]
  real(dp) :: dt_max = 1e-8_dp, epsilon_diel = 1.0_dp
  integer  :: box_size = 8, lvls_max = 6, steps_amr = 2, n_blocks = 2

  type(a5g_t)          :: tree
  type(a5g_mg_t)       :: mg
  type(a5g_ref_info_t) :: ref_info
  real(dp)             :: time, dt, voltage, tbl_x_min, tbl_inv_dx
  real(dp), allocatable :: tbl(:, :)
  integer              :: blk, n, i
  character(len=32)    :: arg

  if (command_argument_count() >= 1) then
     call get_command_argument(1, arg); read(arg, *) lvls_max
  end if
  if (command_argument_count() >= 2) then
     call get_command_argument(2, arg); read(arg, *) n_blocks
  end if
  if (command_argument_count() >= 3) then
     call get_command_argument(3, arg); read(arg, *) applied_fld
  end if
  if (command_argument_count() >= 4) then
     call get_command_argument(4, arg); read(arg, *) domain_len
     seed_r0 = [0.5_dp, 0.5_dp, 0.875_dp] * domain_len
     seed_r1 = seed_r0
  end if

  call init_tree()
  call load_transport_data()
  call st_setup(tree, i_elec, i_pion, i_elec_old, i_pion_old, i_phi, i_fld, i_rhs, &
       -1, i_eps, f_elec, f_fld, i_mobility, i_diffusion, i_alpha, i_eta, .true.)

  ! multigrid: phi, tmp = fld, rhs, eps, n_cycle_base = 8, automatic operators,
  ! sides_bc_pot (src/streamer_3d.f90:693-715) = Neumann sides, grounded
  ! bottom (low z), applied voltage on top
  voltage        = -domain_len * applied_fld
  mg%i_phi       = i_phi
  mg%i_tmp       = i_fld
  mg%i_rhs       = i_rhs
  mg%i_eps       = i_eps
  mg%n_cycle_base = 8
  mg%bc_kind(1:6) = [a5_bc_copy, a5_bc_copy, a5_bc_copy, a5_bc_copy, a5_bc_neg, a5_bc_value]
  mg%bc_val(1:6)  = [0.0_dp, 0.0_dp, 0.0_dp, 0.0_dp, 0.0_dp, voltage]
  mg%auto_ops     = .true.
  call mg3_init_mg(tree, mg)

  time = 0
  dt   = 0

  ! initial conditions and refinement until nothing changes
  ! (src/streamer_3d.f90:58-63)
  do
     call set_init_cond()
     call compute_fld(.true.)
     call adjust(ref_info)
     if (ref_info%n_add == 0 .and. ref_info%n_rm == 0) exit
  end do
  call report(0, 0, 0)

  do blk = 1, n_blocks
     dt = get_max_dt()
     if (dt < 1e-14_dp) then
        print *, "dt getting too small, instability?"
        exit
     end if

     do n = 1, steps_amr
        time = time + dt
        call a3_tree_copy_cc(tree, i_elec, i_elec_old)
        call a3_tree_copy_cc(tree, i_pion, i_pion_old)

        ! two forward Euler steps over dt
        do i = 1, 2
           call st_fluxes_koren(tree, dt)
           call a3_consistent_fluxes(tree, [f_elec])
           call st_update_solution(tree, dt)
           call a3_restrict_tree(tree, i_elec)
           call a3_restrict_tree(tree, i_pion)
           call a3_gc_tree(tree, i_elec, a5_gc_interp_lim, a5_gc_neumann)
           call a3_gc_tree(tree, i_pion, a5_gc_interp_lim, a5_gc_neumann)
           if (i == 1) call compute_fld(.false.)
        end do

        ! explicit trapezoidal rule
        call st_average_dens(tree)
        call compute_fld(.false.)
     end do

     call adjust(ref_info)
     if (ref_info%n_add > 0 .or. ref_info%n_rm > 0) then
        call prolong_to_new_boxes()
        call compute_fld(.false.)
        call a3_tidy_up(tree, 0.9_dp, 0.25_dp, 5000, .false.)
     end if
     call report(blk, ref_info%n_add, ref_info%n_rm)
  end do

  call a3_destroy(tree)

contains

  !> init_tree (src/streamer_3d.f90:150-174): one base box
  subroutine init_tree()
    integer  :: ix_list(3, 1), nb_list(6, 1)
    real(dp) :: dr

    dr = domain_len / box_size
    call a3_init(tree, box_size, n_var_cell, n_var_face, dr, coarsen_to=2, &
         n_boxes=5000, lvls_max=lvls_max)
    ix_list(:, 1) = [1, 1, 1]
    nb_list(:, 1) = -1
    call a3_set_base(tree, ix_list, nb_list)
  end subroutine init_tree

  !> Transport table of the air model: 500 input rows on [0, 3e7] V/m
  !> resampled to 1000 rows (ST_load_transport_data, src/m_streamer.f90:273-333)
  subroutine load_transport_data()
    integer, parameter :: n_in = 500, n_rows = 1000
    real(dp) :: x(n_in), y(n_in)
    integer  :: k

    do k = 1, n_in
       x(k) = 3e7_dp * (k - 1) / (n_in - 1)
    end do
    call st_table_create(tree, 0.0_dp, 3e7_dp, n_rows, 4)
    y = 0.03_dp
    call st_table_set_col(tree, i_mobility, x, y)
    y = 0.2_dp
    call st_table_set_col(tree, i_diffusion, x, y)
    do k = 1, n_in
       if (x(k) > 0) then
          y(k) = 4.3e5_dp * exp(-2e7_dp / x(k))
       else
          y(k) = 0
       end if
    end do
    call st_table_set_col(tree, i_alpha, x, y)
    y = 0
    call st_table_set_col(tree, i_eta, x, y)
    allocate(tbl(n_rows, 4))
    call st_table_get(tree, tbl, tbl_x_min, tbl_inv_dx)
  end subroutine load_transport_data

  !> LT_get_col (fosito/m_lookup_table.f90:100-144) on the host copy
  real(dp) function lt_col(col, x)
    integer, intent(in)  :: col
    real(dp), intent(in) :: x
    real(dp) :: frac, low_frac
    integer  :: low_ix
    frac     = (x - tbl_x_min) * tbl_inv_dx
    low_ix   = ceiling(frac)
    low_frac = low_ix - frac
    if (low_ix < 1) then
       low_ix = 1
       low_frac = 1
    else if (low_ix >= size(tbl, 1)) then
       low_ix = size(tbl, 1) - 1
       low_frac = 0
    end if
    lt_col = low_frac * tbl(low_ix, col) + (1 - low_frac) * tbl(low_ix + 1, col)
  end function lt_col

  !> Distance from r to the segment r0-r1 (src/m_geom.f90:29-46)
  real(dp) function dist_line(r, r0, r1)
    real(dp), intent(in) :: r(3), r0(3), r1(3)
    real(dp) :: len2, along
    len2  = sum((r1 - r0)**2)
    along = sum((r - r0) * (r1 - r0))
    if (along <= 0) then
       dist_line = norm2(r - r0)
    else if (along >= len2) then
       dist_line = norm2(r - r1)
    else
       dist_line = norm2(r - (r0 + along / len2 * (r1 - r0)))
    end if
  end function dist_line

  !> set_init_cond + set_box_eps (src/streamer_3d.f90:217-271) on every box
  subroutine set_init_cond()
    integer, allocatable  :: lvl(:), in_use(:), n_cell(:)
    real(dp), allocatable :: dr(:), r_min(:, :), elec(:, :, :), zero(:, :, :), eps(:, :, :)
    real(dp) :: xyz(3), q
    integer  :: id, i, j, k, nc

    call a5g_box_info(tree, lvl, in_use, n_cell, dr, r_min)
    do id = 1, size(lvl)
       if (in_use(id) == 0) cycle
       nc = n_cell(id)
       allocate(elec(0:nc+1, 0:nc+1, 0:nc+1), zero(0:nc+1, 0:nc+1, 0:nc+1), &
            eps(0:nc+1, 0:nc+1, 0:nc+1))
       do k = 0, nc+1
          do j = 0, nc+1
             do i = 0, nc+1
                xyz = r_min(:, id) + ([i, j, k] - 0.5_dp) * dr(id)
                q   = dist_line(xyz, seed_r0, seed_r1) / seed_width
                elec(i, j, k) = bg_dens + seed_dens * exp(-(q * q))
                if (xyz(1) / domain_len < 0.5_dp .and. xyz(1) / domain_len > 0.375_dp) then
                   eps(i, j, k) = epsilon_diel
                else
                   eps(i, j, k) = 1.0_dp
                end if
             end do
          end do
       end do
       zero = 0
       call a5g_set_cc(tree, id, i_elec, elec)
       call a5g_set_cc(tree, id, i_pion, elec)
       call a5g_set_cc(tree, id, i_phi, zero)
       call a5g_set_cc(tree, id, i_eps, eps)
       deallocate(elec, zero, eps)
    end do
  end subroutine set_init_cond

  !> compute_fld (src/streamer_3d.f90:309-345) with one FMG cycle
  subroutine compute_fld(no_guess)
    logical, intent(in) :: no_guess
    call st_space_charge(tree, UC_elem_charge / UC_eps0)
    call mg3_fas_fmg(tree, mg, .false., no_guess)
    call st_fld_from_pot(tree)
    call a3_gc_tree(tree, i_fld, a5_gc_interp, a5_gc_neumann)
  end subroutine compute_fld

  !> set_ref_flags (src/streamer_3d.f90:177-215) from the per-box maxima,
  !> then a3_adjust_refinement
  subroutine adjust(ri)
    type(a5g_ref_info_t), intent(inout) :: ri
    integer, allocatable  :: lvl(:), in_use(:), n_cell(:), flags(:)
    real(dp), allocatable :: dr(:), r_min(:, :), summ(:, :)
    real(dp) :: crv_phi, max_fld, alpha, adx, boxlen, dist
    integer  :: id

    call a5g_box_info(tree, lvl, in_use, n_cell, dr, r_min)
    call st_box_summaries(tree, summ)
    allocate(flags(size(lvl)))
    flags = a5_kp_ref
    do id = 1, size(lvl)
       if (in_use(id) == 0 .or. lvl(id) < 1) cycle
       crv_phi = summ(1, id)
       max_fld = summ(2, id)
       alpha   = lt_col(i_alpha, max_fld)
       adx     = dr(id) * alpha
       if (adx < 0.1_dp .and. dr(id) < 2.0e-5_dp) flags(id) = a5_rm_ref
       if (adx < 0.1_dp .and. crv_phi < 4.0_dp .and. dr(id) < 5.0e-5_dp) &
            flags(id) = a5_rm_ref
       if (time < 2.5e-9_dp) then
          boxlen = n_cell(id) * dr(id)
          dist = dist_line(r_min(:, id) + 0.5_dp * n_cell(id) * dr(id), seed_r0, seed_r1)
          if (dist - seed_width < boxlen .and. dr(id) > 0.2_dp * seed_width) &
               flags(id) = a5_do_ref
       end if
       if (adx > 1.0_dp .and. crv_phi > 0.1_dp) flags(id) = a5_do_ref
    end do
    call a3_adjust_refinement(tree, flags, ri)
  end subroutine adjust

  !> prolong_to_new_boxes (src/streamer_3d.f90:665-690)
  subroutine prolong_to_new_boxes()
    integer :: lvl
    do lvl = 1, a5g_max_lvl(tree)
       if (lvl > size(ref_info%lvls)) exit
       if (size(ref_info%lvls(lvl)%add) == 0) cycle
       associate (add => ref_info%lvls(lvl)%add)
         call a3_prolong1_to(tree, add, i_elec)
         call a3_prolong1_to(tree, add, i_pion)
         call a3_prolong1_to(tree, add, i_phi)
         call a5g_fill_cc(tree, add, i_eps, 1.0_dp)
         call a3_gc_boxes(tree, add, i_elec, a5_gc_interp_lim, a5_gc_neumann)
         call a3_gc_boxes(tree, add, i_pion, a5_gc_interp_lim, a5_gc_neumann)
         call a3_gc_boxes(tree, add, i_phi, a5_gc_sides_rb, a5_gc_mg_bc)
       end associate
    end do
  end subroutine prolong_to_new_boxes

  !> get_max_dt (src/streamer_3d.f90:274-305)
  real(dp) function get_max_dt()
    real(dp) :: max_fld, min_fld, dr_min, mobility, diff_coeff, alpha
    real(dp) :: dt_cfl, dt_dif, dt_alpha
    call a3_tree_max_cc(tree, i_fld, max_fld)
    call a3_tree_min_cc(tree, i_fld, min_fld)
    dr_min     = a5g_min_dr(tree)
    mobility   = lt_col(i_mobility, max_fld)
    diff_coeff = lt_col(i_diffusion, max_fld)
    alpha      = lt_col(i_alpha, max_fld)
    dt_cfl     = dr_min / (mobility * max_fld)
    dt_dif     = 0.25_dp * dr_min**2 / diff_coeff
    dt_alpha   = 1 / max(mobility * max_fld * alpha, epsilon(1.0_dp))
    get_max_dt = 0.5_dp * min(1/(1/dt_cfl + 1/dt_dif), dt_alpha, dt_max)
  end function get_max_dt

  subroutine report(blk, n_add, n_rm)
    integer, intent(in) :: blk, n_add, n_rm
    real(dp) :: e_max, f_min, f_max
    call a3_tree_max_cc(tree, i_elec, e_max)
    call a3_tree_min_cc(tree, i_fld, f_min)
    call a3_tree_max_cc(tree, i_fld, f_max)
    write(*, "(I0,1X,Z16.16,1X,Z16.16,3(1X,I0),3(1X,Z16.16))") blk, &
         transfer(time, 0_c_int64_t), transfer(dt, 0_c_int64_t), a5g_max_id(tree), &
         n_add, n_rm, transfer(e_max, 0_c_int64_t), transfer(f_min, 0_c_int64_t), &
         transfer(f_max, 0_c_int64_t)
  end subroutine report

end program streamer_3d_a5g
