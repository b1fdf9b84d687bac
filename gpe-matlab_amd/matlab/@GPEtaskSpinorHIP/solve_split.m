function phi = solve_split(task,ddt,niter_inner,niter_outer)
% solve_split - lib/@GPEtaskSpinor/solve_split.m:1-121 on the GPU (gpe_mex).
%
%  Usage :
%    phi = task.solve_split(ddt,niter_inner,niter_outer)
%  phi = {psi1, psi2}.  The coupling is uploaded again whenever a callback changed
%  it (the reference re-reads task.coupling at every outer iteration, lines 61-65).

if(task.omega ~= 0)
    error('GPEtaskSpinorHIP:omega', 'the two-component solver runs without rotation');
end
if(~(task.Ntotal > 0))
    error('GPEtaskSpinorHIP:Ntotal', 'solve_split needs Ntotal > 0 (solve_split.m:30-38)');
end
h = task.hip;
start = task.current_iter;
gpe_mex('omega', h, 0);
gpe_mex('params', h, task.g, task.gamma, task.decay_rate, task.Ntotal, task.mu_init, task.n_recalc, ...
        task.nlincpl);
if(start>0)                                             % solve_split.m:24-28
    phi = task.current_state;
else
    phi = task.init_state;
end
gpe_mex('set_state', h, 0, phi{1});
gpe_mex('set_state', h, 1, phi{2});
VV = task.getVtotal(task.current_time);
gpe_mex('potential', h, 0, VV{1});
gpe_mex('potential', h, 1, VV{2});
coupl = task.coupling;
gpe_mex('coupling', h, coupl);
gpe_mex('split_begin', h, ddt, niter_inner);            % lines 13-57
static = isempty(task.Vtd);
n_rec = task.n_recalc;
dt_outer = ddt*niter_inner;
want_phi = isa(task.user_callback,'function_handle') || task.show_image > 0;
for j=start+1:niter_outer
    if(~isequal(task.coupling, coupl))                  % lines 61-65
        coupl = task.coupling;
        gpe_mex('coupling', h, coupl);
    end
    if(static)
        [time2, muc, ncur] = gpe_mex('split_outer', h, j);
    else
        % the block-end mu applies applyham(phi) at task.current_time (lines 109-114)
        Vc = task.getVtotal(task.current_time);
        gpe_mex('potential_mu', h, 0, Vc{1});
        gpe_mex('potential_mu', h, 1, Vc{2});
        time=(j-1)*dt_outer;
        for jj=1:niter_inner/n_rec
            time2=time+(jj-1)*ddt*n_rec;
            VV = task.getVtotal(time2);
            gpe_mex('potential', h, 0, VV{1});
            gpe_mex('potential', h, 1, VV{2});
            [muc, ncur] = gpe_mex('split_block', h, time2);
        end
    end
    if(want_phi)
        phi = {gpe_mex('get_state', h, 0), gpe_mex('get_state', h, 1)};
    else
        phi = {[], []};
    end
    task.ext_callback(phi,j,time2,muc,ncur);            % line 117
end
if(~static)
    gpe_mex('potential_mu', h, 0, []);
    gpe_mex('potential_mu', h, 1, []);
end
phi = {gpe_mex('get_state', h, 0), gpe_mex('get_state', h, 1)};
task.current_state = phi;
end
